/*
 * spintrack_fields.h -- the user-editable field hooks.
 *
 * The reference has two hooks that users edit and recompile to get fields
 * its parameter file cannot describe: pulse() and grad() in
 * src/integrator.cpp:40-46 (README.md, "Modification of Field Gradients and
 * Magnetic Fields").  Both are compiled for the CPU and the GPU from one
 * source.  This header is the same two hooks for this library: edit them here
 * (or keep them in a file of your own and build with
 * `make -C spin-tracking_amd FIELDS=/path/hooks.h`, which defines
 * ST_FIELDS_FILE), rebuild, and every kernel evaluates the new field.  The
 * CPU oracle (oracle/, test infrastructure) includes the same header, so its
 * parity tests check the kernels against the edited fields too
 * (`make -C oracle FIELDS=/path/hooks.h`).
 *
 * The hooks are plain C (compiled by gcc for the oracle, and by hipcc for
 * the host and the device).  They see the same arguments as the reference's:
 *
 *   st_field_pulse(t, a, w, axis, B)  pulse(t, a, w) of src/integrator.cpp:40-42:
 *                                     the oscillating field at time t (tesla),
 *                                     a and w from the options, axis from
 *                                     st_ext.pulse_axis (the reference's hook
 *                                     writes it on z);
 *   st_field_grad(r, G, B)            grad(pos, opt) of src/integrator.cpp:44-46:
 *                                     the gradient field at position r (tesla),
 *                                     G the rows Gx, Gy, Gz of the options
 *                                     (G[3i+j] = dB_i/dx_j).
 *
 * The total field is B = pulse + B0 + (1/c^2) v x E + grad, added in that
 * order (findCrossTerm, src/integrator.cpp:63-70).
 *
 * Two traits tell the library what it may assume about the hooks.  Set them
 * to 0 when an edit breaks them:
 *
 *   ST_FIELDS_GRAD_LINEAR     grad is the linear map G applied to r.  Along a
 *                             ballistic segment the field is then a quadratic
 *                             polynomial in time; the fast variant
 *                             (ST_VARIANT_FAST) integrates that polynomial, and
 *                             the adjoint's gradient with respect to G is
 *                             defined through it.
 *   ST_FIELDS_PULSE_COS_AXIS  pulse is a*cos(w*t) on one axis (zero when
 *                             a == 0, so the kernels may skip it).
 *
 * With either trait 0, ST_VARIANT_EXACT runs the hooks as written and
 * st_create refuses ST_VARIANT_FAST, st_adjoint and integratorType 6's fast
 * sampler with ST_E_UNSUPPORTED.
 */
#ifndef SPINTRACK_FIELDS_H
#define SPINTRACK_FIELDS_H

#include <math.h>

#if defined(__HIPCC__)
#define ST_FIELD_FN static inline __host__ __device__
#else
#define ST_FIELD_FN static inline
#endif

#ifdef ST_FIELDS_FILE
#include ST_FIELDS_FILE
#else

#define ST_FIELDS_GRAD_LINEAR 1
#define ST_FIELDS_PULSE_COS_AXIS 1

/* src/integrator.cpp:40-42: {0, 0, a*cos(w*t)}, here on the chosen axis */
ST_FIELD_FN void st_field_pulse(double t, double a, double w, int axis, double B[3]) {
    const double p = a * cos(w * t);
    B[0] = axis == 0 ? p : 0.0;
    B[1] = axis == 1 ? p : 0.0;
    B[2] = axis == 2 ? p : 0.0;
}

/* src/integrator.cpp:44-46: {dot(pos, Gx), dot(pos, Gy), dot(pos, Gz)}, each
 * dot product summed left to right (src/double3.cpp:93-95) */
ST_FIELD_FN void st_field_grad(const double r[3], const double G[9], double B[3]) {
    B[0] = r[0] * G[0] + r[1] * G[1] + r[2] * G[2];
    B[1] = r[0] * G[3] + r[1] * G[4] + r[2] * G[5];
    B[2] = r[0] * G[6] + r[1] * G[7] + r[2] * G[8];
}

#endif /* ST_FIELDS_FILE */

#if !defined(ST_FIELDS_GRAD_LINEAR) || !defined(ST_FIELDS_PULSE_COS_AXIS)
#error "a field-hook file must define ST_FIELDS_GRAD_LINEAR and ST_FIELDS_PULSE_COS_AXIS"
#endif

/* 1 when the fast variant's per-segment field polynomial is exact for the hooks */
#define ST_FIELDS_POLYNOMIAL (ST_FIELDS_GRAD_LINEAR && ST_FIELDS_PULSE_COS_AXIS)

#endif /* SPINTRACK_FIELDS_H */
