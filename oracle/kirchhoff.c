/* Kirchhoff migration of one shot -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Sequential C restatement of
 *   oracle_eikonal          solve_eikonal_FSM (solve_eikonal_FSM.f90:1-83): T = 1e9,
 *                           T(source) = 0, four Gauss-Seidel sweeps in the reference's loop
 *                           order (j outer, i inner; :30-58), local_solver (:63-79);
 *   oracle_kirchhoff_image  the imaging loop of calculate_PSDM_kirchoff.f90:394-447
 *                           (receiver, column, row loops; acosf / expf of the host libm, as
 *                           the reference's compiled Fortran calls them).
 * FP32 throughout, every operation rounded separately (built with -ffp-contract=off).
 * Pinned by tests/test_kirchhoff_oracle.py: the eikonal bit for bit against the reference's
 * solve_eikonal_FSM compiled into oracle/_ref; the imaging loop by special cases and
 * invariants (it is inline in the reference's driver, not a callable routine).
 */
#include <math.h>

static float local_solver(float a, float b, float dx, float dy, float v_inv) {
    const float aa = 1.0f / (dx * dx) + 1.0f / (dy * dy);
    const float bb = -2.0f * (a / (dx * dx) + b / (dy * dy));
    const float cc = (a * a) / (dx * dx) + (b * b) / (dy * dy) - v_inv * v_inv;
    const float disc = bb * bb - 4.0f * aa * cc;
    float t;
    if (disc >= 0.0f) {
        t = (-bb + sqrtf(disc)) / (2.0f * aa);
        const float m = a > b ? a : b;
        if (t < m) t = m + v_inv * (dx < dy ? dx : dy);
    } else {
        t = (a < b ? a : b) + v_inv * (dx < dy ? dx : dy);
    }
    return t;
}

static float fmin2(float x, float y) { return x < y ? x : y; }

/* v, T: (ny, nx) column-major; sx, sz: 1-based column and row of the source */
void oracle_eikonal(const float *v, int ny, int nx, float dx, float dy, int sx, int sz, float *T) {
    for (long e = 0; e < (long)nx * ny; ++e) T[e] = 1.0e9f;
    T[(long)(sx - 1) * ny + (sz - 1)] = 0.0f;
#define AT(j, i) T[(long)((i) - 1) * ny + ((j) - 1)]
    for (int sweep = 1; sweep <= 4; ++sweep) {
        int is, ie, di, js, je, dj;
        switch (sweep) {
        case 1: is = 1; ie = nx; di = 1; js = 1; je = ny; dj = 1; break;
        case 2: is = nx; ie = 1; di = -1; js = 1; je = ny; dj = 1; break;
        case 3: is = nx; ie = 1; di = -1; js = ny; je = 1; dj = -1; break;
        default: is = 1; ie = nx; di = 1; js = ny; je = 1; dj = -1; break;
        }
        for (int j = js; dj > 0 ? j <= je : j >= je; j += dj)
            for (int i = is; di > 0 ? i <= ie : i >= ie; i += di) {
                if (AT(j, i) == 0.0f) continue;
                const float a = fmin2(AT(j, i > 1 ? i - 1 : 1), AT(j, i < nx ? i + 1 : nx));
                const float b = fmin2(AT(j > 1 ? j - 1 : 1, i), AT(j < ny ? j + 1 : ny, i));
                const float v_inv = 1.0f / v[(long)(i - 1) * ny + (j - 1)];
                const float t = local_solver(a, b, dx, dy, v_inv);
                if (t < AT(j, i)) AT(j, i) = t;
            }
    }
#undef AT
}

/* tts (ny, nx), ttr (ny, nx, nrec) column-major; data (nrec, ld) trace-major; img (ny, nx) */
void oracle_kirchhoff_image(const float *tts, const float *ttr, const float *data, int ld, int nrec, int nt_prop,
                            int ny, int nx, float dd, float dt, float vp0, float theta_max, float theta_taper,
                            float wmax, float *img) {
    const float tmin = 0.02f, eps = 1.0e-6f, pi = 3.14159265f;
    const long n = (long)nx * ny;
    for (long p = 0; p < n; ++p) img[p] = 0.0f;
    for (int k = 0; k < nrec; ++k)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                const long p = (long)i * ny + j;
                const float z_img = (float)j * dd;
                const float tt_tot = tts[p] + ttr[(long)k * n + p];
                const int it = 1 + (int)floorf(tt_tot / dt);
                if (it <= 1 || it >= nt_prop) continue;
                const float rs = vp0 * (tts[p] > tmin ? tts[p] : tmin);
                const float rr = vp0 * (ttr[(long)k * n + p] > tmin ? ttr[(long)k * n + p] : tmin);
                float cos_theta = z_img / (rs + rr + 1.e-6f);
                cos_theta = cos_theta < 1.0f ? cos_theta : 1.0f;
                cos_theta = cos_theta > 0.0f ? cos_theta : 0.0f;
                const float theta = acosf(cos_theta) * 180.0f / pi;
                float w_theta;
                if (theta <= theta_max) {
                    w_theta = 1.0f;
                } else {
                    const float u = (theta - theta_max) / theta_taper;
                    w_theta = expf(-(u * u));
                }
                const float q = cos_theta / sqrtf(rs * rr + eps);
                const float weight = w_theta * (q < wmax ? q : wmax);
                const float amp = (data[(long)k * ld + it] - data[(long)k * ld + it - 2]) / (2.0f * dt);
                img[p] = img[p] + weight * amp;
            }
}
