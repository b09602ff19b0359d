// Host check of include/fvm2d_frame.hpp (built and run by
// tests/test_record_io.py): writes the w x h test image
// gray(i, j) = (7 i + 13 j) mod 256 as a record-mode frame.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fvm2d_frame.hpp"

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s out.png w h\n", argv[0]);
        return 2;
    }
    const int w = std::atoi(argv[2]), h = std::atoi(argv[3]);
    std::vector<unsigned char> gray((size_t)w * h);
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++) gray[(size_t)j * w + i] = (unsigned char)((7 * i + 13 * j) % 256);
    fvm2d::write_frame_png(argv[1], w, h, gray.data());
    return 0;
}
