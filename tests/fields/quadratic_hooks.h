/*
 * Field hooks for the hook tests (built into spin-tracking_amd/libspintrack_fieldtest.so
 * and oracle/liboracle_fieldtest.so by __graft_entry__.build()): a gradient
 * with a quadratic term and a pulse that rotates in the x-y plane.  Neither
 * fits the fast variant's per-segment polynomial, so both traits are 0.
 */
#define ST_FIELDS_GRAD_LINEAR 0
#define ST_FIELDS_PULSE_COS_AXIS 0

/* the linear gradient of the options plus Q*(z^2 - x^2, x*y, x*z), Q in T/m^2 */
#define ST_TEST_Q 2.0e-6

ST_FIELD_FN void st_field_grad(const double r[3], const double G[9], double B[3]) {
    B[0] = (r[0] * G[0] + r[1] * G[1] + r[2] * G[2]) + ST_TEST_Q * (r[2] * r[2] - r[0] * r[0]);
    B[1] = (r[0] * G[3] + r[1] * G[4] + r[2] * G[5]) + ST_TEST_Q * (r[0] * r[1]);
    B[2] = (r[0] * G[6] + r[1] * G[7] + r[2] * G[8]) + ST_TEST_Q * (r[0] * r[2]);
}

/* a field of size a rotating about the pulse axis at angular frequency w */
ST_FIELD_FN void st_field_pulse(double t, double a, double w, int axis, double B[3]) {
    const double c = a * cos(w * t), s = a * sin(w * t);
    B[0] = axis == 0 ? 0.0 : c;
    B[1] = axis == 1 ? 0.0 : (axis == 0 ? c : s);
    B[2] = axis == 2 ? 0.0 : s;
}
