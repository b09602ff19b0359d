/*
 * fvm2d measurement probes: C ABI of libfvm2d_probes.so.
 *
 * Diagnostics with no counterpart in the reference, kept out of the solver
 * library libfvm2d.so.  They measure the card the solver runs on: the HBM
 * byte counters' calibration, the FP64 VALU issue rate and the accuracy of
 * the hardware reciprocal estimates the FAST arithmetic refines.  Every
 * function returns 0 or a negative FVM2D_PROBE_ERR_* code; the message of the
 * last failure on the calling thread is fvm2d_probe_last_error().
 */
#ifndef FVM2D_PROBES_H
#define FVM2D_PROBES_H

#ifdef __cplusplus
extern "C" {
#endif

#define FVM2D_PROBE_ERR_ARG (-1)
#define FVM2D_PROBE_ERR_HIP (-2)
#define FVM2D_PROBE_ERR_NODEVICE (-4)

const char *fvm2d_probe_last_error(void);

/* reps launches of a plain device-to-device copy of n_doubles FP64 values
 * (16-byte accesses, two buffers allocated and freed by the call).  Its
 * FETCH_SIZE / WRITE_SIZE counters calibrate the step kernel's measured HBM
 * traffic.  avg_ms_out: mean duration of one copy. */
int fvm2d_hbm_copy_probe(long long n_doubles, int reps, double *avg_ms_out);

/* Sustained wave-level instruction rate per SIMD of one VALU kind with 16
 * independent chains per lane on every SIMD (kind 0 v_fma_f64, 1 v_mul_f64,
 * 2 v_add_f64, 3 v_rcp_f64, 4 32-bit integer ops). */
int fvm2d_valu_probe(int kind, double *per_simd_per_s);

/* Maximum relative error, over 2^18 arguments 2^u (u uniform in [-8, 8)),
 * of the device's raw v_rcp_f64, the FAST path's refined reciprocal, raw
 * v_rsq_f64 and the refined reciprocal square root (out[0..3]), against
 * long-double host references. */
int fvm2d_recip_probe(double *max_rel_err4);

#ifdef __cplusplus
}
#endif
#endif /* FVM2D_PROBES_H */
