/*
 * vfs_oracle.c -- float64 C restatement of the VFS coarse-to-fine solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a second, compiled form of the oracle in
 * oracle/vfs_oracle.py: the same kernels, the same sampler semantics and the same
 * host loop, pixel by pixel in double precision, single-threaded.  It exists so
 * that the CPU baseline of bench.py can time whole solves of the 800 x 800
 * workload in seconds rather than minutes, and so that the NumPy restatement has
 * an independent implementation to agree with (tests/test_oracle_c.py).  Only
 * tests/ and the cpu_baseline leg of bench.py load it; the product
 * (vfs-python_amd/) never does.
 *
 * Each function follows the kernel of the reference it names (file:line in
 * eshibusawa/VFS-Python) and the operation order of its NumPy restatement, so the
 * two agree to rounding:
 *   sampler          Python/vfs.cu:35-79 (bilinear, mirror addressing)
 *   pyramids         Python/vfs.py:93-218, Python/vfs.cu:145-184
 *   warp             Python/vfs.cu:186-202, :283-302
 *   Gaussian         Python/vfs.cu:81-102, :204-226
 *   tensor, eta      Python/vfs.cu:104-118, :228-281
 *   derivatives      Python/vfs.cu:120-142, :304-324
 *   dual             Python/vfs.cu:326-419
 *   threshold        Python/vfs.cu:421-505
 *   primal           Python/vfs.cu:507-581
 *   median           Python/vfs.cu:583-639
 *   clamp            Python/vfs.cu:641-676
 *   upsampling       Python/vfs.cu:69-79, :678-705
 *   solve loop       Python/vfs.py:460-533
 * The tau / sigma / mu schedule is an input (the NumPy restatement's pd_schedule,
 * which reproduces Python/vfs.py:494-515 under NumPy's scalar rules).
 *
 * Build: gcc -O3 -ffp-contract=off -fopenmp -std=c99 -fPIC -shared vfs_oracle.c -o libvfs_oracle.so -lm
 * (no -ffast-math, no contraction: the restatement keeps IEEE double semantics).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ETA_P 3.0
#define ETA_Q 2.0
#define MIN_NORMALIZED_GRAD 1e-8
#define MIN_TENSOR_VALUE 1e-8
#define EPS 1e-7

typedef struct vfs_oracle_params {
    double beta, gamma, alpha0, alpha1, lambda, f_scale, limit_range;
    int32_t n_level, n_warp, n_iter, median_k;
    int32_t n_threads; /* OpenMP threads over rows (every pixel is computed alone: any count gives the same bits) */
} vfs_oracle_params;

static int g_threads = 1;

/* half-sample symmetric reflection with period 2n */
static long mirror_index(long i, long n)
{
    long p = 2 * n;
    long m = i % p;
    if (m < 0) m += p;
    return m >= n ? p - 1 - m : m;
}

static long clampi(long i, long n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

/* bilinear sample of channel ch of an (H, W, C) image at texel coordinates (x, y) */
static double sample_linear(const double* img, long H, long W, long C, long ch, double x, double y)
{
    double xc = isnan(x) ? 0.0 : x, yc = isnan(y) ? 0.0 : y;
    xc = xc < -1e7 ? -1e7 : (xc > 1e7 ? 1e7 : xc);
    yc = yc < -1e7 ? -1e7 : (yc > 1e7 ? 1e7 : yc);
    const double x0 = floor(xc), y0 = floor(yc);
    const double ax = x - x0, ay = y - y0;
    const long ix0 = mirror_index((long)x0, W), ix1 = mirror_index((long)x0 + 1, W);
    const long iy0 = mirror_index((long)y0, H), iy1 = mirror_index((long)y0 + 1, H);
    const double t00 = img[(iy0 * W + ix0) * C + ch], t10 = img[(iy0 * W + ix1) * C + ch];
    const double t01 = img[(iy1 * W + ix0) * C + ch], t11 = img[(iy1 * W + ix1) * C + ch];
    const double r0 = t00 + ax * (t10 - t00);
    const double r1 = t01 + ax * (t11 - t01);
    return r0 + ay * (r1 - r0);
}

/* downSamplingImage / downSamplingVectorField: four linear taps at +-1/4 output pixel */
static void downsample_bilinear(double* out, long Hd, long Wd, const double* src, long Hs, long Ws, long C)
{
    const double rx = (double)Ws / (double)Wd, ry = (double)Hs / (double)Hd;
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < Hd; ++Y)
        for (long X = 0; X < Wd; ++X) {
            const double xc = (X + 0.5) * rx - 0.5, yc = (Y + 0.5) * ry - 0.5;
            const double xm = (X + 0.25) * rx - 0.5, xp = (X + 0.75) * rx - 0.5;
            const double ym = (Y + 0.25) * ry - 0.5, yp = (Y + 0.75) * ry - 0.5;
            for (long c = 0; c < C; ++c)
                out[(Y * Wd + X) * C + c] =
                    0.25 * (sample_linear(src, Hs, Ws, C, c, xm, yc) + sample_linear(src, Hs, Ws, C, c, xp, yc) +
                            sample_linear(src, Hs, Ws, C, c, xc, ym) + sample_linear(src, Hs, Ws, C, c, xc, yp));
        }
}

/* downSamplingMask: texel ((2X+1) Ws) div (2 Wd) */
static void downsample_mask(uint8_t* out, long Hd, long Wd, const uint8_t* src, long Hs, long Ws)
{
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < Hd; ++Y)
        for (long X = 0; X < Wd; ++X)
            out[Y * Wd + X] = src[((2 * Y + 1) * Hs) / (2 * Hd) * Ws + ((2 * X + 1) * Ws) / (2 * Wd)];
}

/* warpingImage / findWarpingVector at equal sizes: S_lin(src, X + w.x, Y + w.y) */
static void warping_image(double* out, const double* src, long C, const double* warp, long H, long W)
{
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const long i = Y * W + X;
            const double xs = X + warp[2 * i], ys = Y + warp[2 * i + 1];
            for (long c = 0; c < C; ++c) out[i * C + c] = sample_linear(src, H, W, C, c, xs, ys);
        }
}

#define NB(a, X, Y, dx, dy) (a)[clampi((Y) + (dy), H) * W + clampi((X) + (dx), W)]

/* gaussianMasked: binomial 3 x 3 / 16, 0 where the mask is 0 */
static void gaussian_masked(double* out, const double* img, const uint8_t* mask, long H, long W)
{
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const double v0 = img[Y * W + X];
            const double v1 = NB(img, X, Y, -1, -1) + NB(img, X, Y, -1, 1) + NB(img, X, Y, 1, -1) + NB(img, X, Y, 1, 1);
            const double v2 = NB(img, X, Y, -1, 0) + NB(img, X, Y, 1, 0) + NB(img, X, Y, 0, -1) + NB(img, X, Y, 0, 1);
            const double o = 0.25 * v0 + 0.0625 * v1 + 0.125 * v2;
            out[Y * W + X] = mask[Y * W + X] == 0 ? 0.0 : o;
        }
}

/* computeTensorEtaMasked with computeEta */
static void compute_tensor_eta(double* abc, double* eta, const double* img, const uint8_t* mask, long H, long W,
                               const vfs_oracle_params* p)
{
    const double a11 = p->alpha1 * p->alpha1, a004 = 4.0 * p->alpha0 * p->alpha0;
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const long i = Y * W + X;
            if (mask[i] == 0) {
                for (int k = 0; k < 3; ++k) abc[3 * i + k] = eta[3 * i + k] = 0.0;
                continue;
            }
            const double i11 = NB(img, X, Y, 1, 1);
            const double gx = i11 - NB(img, X, Y, 0, 1);
            const double gy = i11 - NB(img, X, Y, 1, 0);
            const double ng = hypot(gx, gy);
            const int small = ng < MIN_NORMALIZED_GRAD;
            const double nx = small ? 1.0 : gx / ng, ny = small ? 0.0 : gy / ng;
            const double tx = ny, ty = -nx;
            double w = exp(-p->beta * pow(ng, p->gamma));
            w = w > MIN_TENSOR_VALUE ? w : MIN_TENSOR_VALUE;
            const double a = w * nx * nx + tx * tx;
            const double b = w * ny * ny + ty * ty;
            const double c = w * nx * ny + tx * ty;
            abc[3 * i] = a;
            abc[3 * i + 1] = b;
            abc[3 * i + 2] = c;
            eta[3 * i] = a11 * (a * a + b * b + 2.0 * c * c) + a11 * ((a + c) * (a + c) + (b + c) * (b + c));
            eta[3 * i + 1] = a11 * (b * b + c * c) + a004;
            eta[3 * i + 2] = a11 * (a * a + c * c) + a004;
        }
}

static double taps(const double* img, long H, long W, double X, double Y, double ex, double ey)
{
    double v = sample_linear(img, H, W, 1, 0, X - 2 * ex, Y - 2 * ey);
    v = v - 8.0 * sample_linear(img, H, W, 1, 0, X - ex, Y - ey);
    v = v + 8.0 * sample_linear(img, H, W, 1, 0, X + ex, Y + ey);
    v = v - sample_linear(img, H, W, 1, 0, X + 2 * ex, Y + 2 * ey);
    return v;
}

/* computeDerivatives: (Iu, Iz) */
static void compute_derivatives(double* deriv, const double* ow, const double* ref, const double* tvf, long H, long W)
{
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const long i = Y * W + X;
            const double nrm = hypot(tvf[2 * i], tvf[2 * i + 1]);
            const double ex = tvf[2 * i] / nrm, ey = tvf[2 * i + 1] / nrm;
            deriv[2 * i] = (taps(ref, H, W, X, Y, ex, ey) + taps(ow, H, W, X, Y, ex, ey)) / 24.0;
            deriv[2 * i + 1] = ow[i] - ref[i];
        }
}

/* updateDualVariables: new p, q into pn_buf, qn_buf; tp written */
static void update_dual(const double* p, double* tp, const double* q, double sigma, const double* u_,
                        const double* v_, const double* abc, const uint8_t* m, long H, long W,
                        const vfs_oracle_params* prm, double* pn_buf, double* qn_buf)
{
    const double sp = prm->alpha1 * sigma / ETA_P, sq = prm->alpha0 * sigma / ETA_Q;
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const long i = Y * W + X;
            if (m[i] == 0) {
                pn_buf[2 * i] = pn_buf[2 * i + 1] = 0.0;
                tp[2 * i] = tp[2 * i + 1] = 0.0;
                for (int k = 0; k < 4; ++k) qn_buf[4 * i + k] = 0.0;
                continue;
            }
            const int mE = NB(m, X, Y, 1, 0) != 0, mS = NB(m, X, Y, 0, 1) != 0;
            const long iE = Y * W + clampi(X + 1, W), iW = Y * W + clampi(X - 1, W);
            const long iS = clampi(Y + 1, H) * W + X, iN = clampi(Y - 1, H) * W + X;
            const double u = u_[i], vx = v_[2 * i], vy = v_[2 * i + 1];
            const double gux = mE ? u_[iE] - u - vx : u - u_[iW] - vx;
            const double gvx = mE ? v_[2 * iE] - vx : vx - v_[2 * iW];
            const double gvw = mE ? v_[2 * iE + 1] - vy : vy - v_[2 * iW + 1];
            const double guy = mS ? u_[iS] - u - vy : u - u_[iN] - vy;
            const double gvy = mS ? v_[2 * iS + 1] - vy : vy - v_[2 * iN + 1];
            const double gvz = mS ? v_[2 * iS] - vx : vx - v_[2 * iN];
            const double a = abc[3 * i], b = abc[3 * i + 1], c = abc[3 * i + 2];
            const double tux = a * gux + c * guy, tuy = c * gux + b * guy;
            double px = p[2 * i] + sp * tux, py = p[2 * i + 1] + sp * tuy;
            double n = hypot(px, py);
            n = n > 1.0 ? n : 1.0;
            px = px / n;
            py = py / n;
            pn_buf[2 * i] = px;
            pn_buf[2 * i + 1] = py;
            tp[2 * i] = a * px + c * py;
            tp[2 * i + 1] = c * px + b * py;
            const double q0 = q[4 * i] + sq * gvx, q1 = q[4 * i + 1] + sq * gvy, q2 = q[4 * i + 2] + sq * gvz,
                         q3 = q[4 * i + 3] + sq * gvw;
            double nq = sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
            nq = nq > 1.0 ? nq : 1.0;
            qn_buf[4 * i] = q0 / nq;
            qn_buf[4 * i + 1] = q1 / nq;
            qn_buf[4 * i + 2] = q2 / nq;
            qn_buf[4 * i + 3] = q3 / nq;
        }
}

/* l1Thresholding: u_out may not alias u (the NumPy restatement returns a new array) */
static void l1_thresholding(double* out, double tau, const double* tp, const double* u, const double* u_,
                            const double* deriv, const double* eta, const uint8_t* m, long H, long W, double lam)
{
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const long i = Y * W + X;
            if (m[i] == 0) {
                out[i] = 0.0;
                continue;
            }
            const int mE = NB(m, X, Y, 1, 0) != 0, mW = NB(m, X, Y, -1, 0) != 0;
            const int mS = NB(m, X, Y, 0, 1) != 0, mN = NB(m, X, Y, 0, -1) != 0;
            const long iW = Y * W + clampi(X - 1, W), iN = clampi(Y - 1, H) * W + X;
            const double tpx = tp[2 * i], tpy = tp[2 * i + 1], tpWx = tp[2 * iW], tpNy = tp[2 * iN + 1];
            const double gx = (mE && mW) ? tpx - tpWx : (!mE ? -tpWx : tpx);
            const double gy = (mS && mN) ? tpy - tpNy : (!mS ? -tpNy : tpy);
            const double iu = deriv[2 * i], iz = deriv[2 * i + 1], e0 = eta[3 * i];
            const double te = fabs(e0) < EPS ? tau : tau / e0;
            const double ut = u_[i] + te * (gx + gy);
            const double rho = iu * (ut - u[i]) + iz;
            const double bound = lam * te * iu * iu;
            const double d0 = ut - u[i];
            const double du_in = fabs(iu) < EPS ? d0 : d0 - rho / iu;
            double du;
            if (rho <= bound && rho >= -bound)
                du = du_in;
            else if (rho < -bound)
                du = d0 + lam * te * iu;
            else if (rho > bound)
                du = d0 - lam * te * iu;
            else
                du = 0.0;
            out[i] = u[i] + du;
        }
}

/* updatePrimalVariables: u_bar, v_bar in place (each pixel reads only its own values) */
static void update_primal(double tau, double mu, const double* p, const double* q, const double* u, double* u_,
                          double* v_, const double* abc, const double* eta, const uint8_t* m, long H, long W,
                          const vfs_oracle_params* prm)
{
    const double a0 = prm->alpha0, a1 = prm->alpha1;
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const long i = Y * W + X;
            if (m[i] == 0) {
                u_[i] = 0.0;
                v_[2 * i] = v_[2 * i + 1] = 0.0;
                continue;
            }
            const int mE = NB(m, X, Y, 1, 0) != 0, mW = NB(m, X, Y, -1, 0) != 0;
            const int mS = NB(m, X, Y, 0, 1) != 0, mN = NB(m, X, Y, 0, -1) != 0;
            const long iW = Y * W + clampi(X - 1, W), iN = clampi(Y - 1, H) * W + X;
            const double gqx = (mE ? q[4 * i] : 0.0) - (mW ? q[4 * iW] : 0.0);
            const double gqw = (mE ? q[4 * i + 3] : 0.0) - (mW ? q[4 * iW + 3] : 0.0);
            const double gqy = (mS ? q[4 * i + 1] : 0.0) - (mN ? q[4 * iN + 1] : 0.0);
            const double gqz = (mS ? q[4 * i + 2] : 0.0) - (mN ? q[4 * iN + 2] : 0.0);
            const double a = abc[3 * i], b = abc[3 * i + 1], c = abc[3 * i + 2];
            const double tux = a * p[2 * i] + c * p[2 * i + 1];
            const double tuy = c * p[2 * i] + b * p[2 * i + 1];
            const double e1 = eta[3 * i + 1], e2 = eta[3 * i + 2];
            const double t1 = fabs(e1) < EPS ? tau : tau / e1;
            const double t2 = fabs(e2) < EPS ? tau : tau / e2;
            const double vx = t1 * (a1 * tux + a0 * (gqx + gqz)) + v_[2 * i];
            const double vy = t2 * (a1 * tuy + a0 * (gqw + gqy)) + v_[2 * i + 1];
            const double un = mu * (u[i] - u_[i]) + u[i];
            const double vxn = mu * (vx - v_[2 * i]) + vx;
            const double vyn = mu * (vy - v_[2 * i + 1]) + vy;
            u_[i] = un;
            v_[2 * i] = vxn;
            v_[2 * i + 1] = vyn;
        }
}

/* medianFilter: k x k clamp-to-edge window, element k*k/2 of the window sorted by
 * insertion (the order statistic, whatever the sort); masked pixels copy */
static void median_filter(double* out, const double* u, const uint8_t* m, long H, long W, int k)
{
    const int h = k / 2, n = k * k;
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < H; ++Y)
        for (long X = 0; X < W; ++X) {
            const long i = Y * W + X;
            if (m[i] == 0) {
                out[i] = u[i];
                continue;
            }
            double win[49];
            int c = 0;
            for (int dy = -h; dy <= h; ++dy)
                for (int dx = -h; dx <= h; ++dx) win[c++] = u[clampi(Y + dy, H) * W + clampi(X + dx, W)];
            for (int a = 1; a < n; ++a) {
                const double v = win[a];
                int b = a - 1;
                while (b >= 0 && win[b] > v) {
                    win[b + 1] = win[b];
                    --b;
                }
                win[b + 1] = v;
            }
            out[i] = win[n / 2];
        }
}

/* computeOpticalFlowWithClamp: flow, u0 in place */
static void flow_with_clamp(double* flow, double* u0, const double* uf, const double* tv, const uint8_t* m, long n,
                            double limit)
{
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long i = 0; i < n; ++i) {
        if (m[i] == 0) continue;
        double d = uf[i] - u0[i];
        d = d < limit ? d : limit;
        u0[i] = u0[i] + d;
        flow[2 * i] = flow[2 * i] + d * tv[2 * i];
        flow[2 * i + 1] = flow[2 * i + 1] + d * tv[2 * i + 1];
    }
}

/* upSamplingVectorField(1): f_scale * S_lin(src, (X + 0.5) Ws / Wd - 0.5) */
static void upsampling(double* out, long Hd, long Wd, const double* src, long Hs, long Ws, long C, double f_scale)
{
    const double rx = (double)Ws / (double)Wd, ry = (double)Hs / (double)Hd;
    #pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long Y = 0; Y < Hd; ++Y)
        for (long X = 0; X < Wd; ++X) {
            const double xs = (X + 0.5) * rx - 0.5, ys = (Y + 0.5) * ry - 0.5;
            for (long c = 0; c < C; ++c) out[(Y * Wd + X) * C + c] = f_scale * sample_linear(src, Hs, Ws, C, c, xs, ys);
        }
}

/* VFS.solve_stereo (Python/vfs.py:460-533) with the pyramids of Python/vfs.py:121-218.
 * img_ref, img_other, mask (H, W) uint8; tvf, cvf (H, W, 2) float32; taus, sigmas, mus:
 * n_iter schedule values.  flow_out (H, W, 2) double.  Returns 0, or -1 when memory runs
 * out, -22 for bad arguments. */
int vfs_oracle_solve(const uint8_t* img_ref, const uint8_t* img_other, const uint8_t* mask, const float* tvf,
                     const float* cvf, int H0, int W0, const vfs_oracle_params* prm, const double* taus,
                     const double* sigmas, const double* mus, double* flow_out)
{
    if (!img_ref || !img_other || !mask || !tvf || !cvf || !prm || !flow_out || H0 <= 0 || W0 <= 0) return -22;
    const int L = prm->n_level, NW = prm->n_warp, NI = prm->n_iter;
    if (L < 1 || L > 64 || NW < 0 || NI < 0 || (NI > 0 && (!taus || !sigmas || !mus))) return -22;
    if (prm->median_k != 3 && prm->median_k != 5 && prm->median_k != 7) return -22;
    g_threads = prm->n_threads > 0 ? prm->n_threads : 1;
    long hs[64], ws[64];
    {
        long h = H0, w = W0;
        for (int k = 0; k < L; ++k) {
            hs[k] = h;
            ws[k] = w;
            h = (long)((double)h / prm->f_scale);
            w = (long)((double)w / prm->f_scale);
            if (k + 1 < L && (h <= 0 || w <= 0)) return -22;
        }
    }
    const long n0 = (long)H0 * W0;
    int rc = -1;
    uint8_t* masks[64] = {0};
    double *trs[64] = {0}, *refs[64] = {0}, *others[64] = {0};
    double *warp = 0, *u = 0, *ub = 0, *ref_s = 0, *abc = 0, *eta = 0, *tv = 0, *ow = 0, *deriv = 0, *p = 0, *q = 0,
           *tp = 0, *vb = 0, *u0 = 0, *uf = 0, *pnb = 0, *qnb = 0, *tmp = 0, *tmp2 = 0;
#define ALLOC(ptr, count)                                  \
    do {                                                   \
        (ptr) = malloc(sizeof(*(ptr)) * (size_t)(count));  \
        if (!(ptr)) goto done;                             \
    } while (0)
    for (int k = 0; k < L; ++k) {
        const long n = hs[k] * ws[k];
        ALLOC(masks[k], n);
        ALLOC(trs[k], 2 * n);
        ALLOC(refs[k], n);
        ALLOC(others[k], n);
    }
    ALLOC(warp, 2 * n0);
    ALLOC(u, 2 * n0); /* swapped with tmp, which also holds two-channel fields */
    ALLOC(ub, n0);
    ALLOC(ref_s, n0);
    ALLOC(abc, 3 * n0);
    ALLOC(eta, 3 * n0);
    ALLOC(tv, 2 * n0);
    ALLOC(ow, n0);
    ALLOC(deriv, 2 * n0);
    ALLOC(p, 2 * n0);
    ALLOC(q, 4 * n0);
    ALLOC(tp, 2 * n0);
    ALLOC(vb, 2 * n0);
    ALLOC(u0, n0);
    ALLOC(uf, n0);
    ALLOC(pnb, 2 * n0);
    ALLOC(qnb, 4 * n0);
    ALLOC(tmp, 2 * n0);
    ALLOC(tmp2, 2 * n0);
#undef ALLOC
    /* pyramids: masks, translation field, images (the other view de-rotated at level 0) */
    memcpy(masks[0], mask, (size_t)n0);
    for (long i = 0; i < 2 * n0; ++i) trs[0][i] = (double)tvf[i];
    for (long i = 0; i < n0; ++i) refs[0][i] = (double)img_ref[i] / 256.0;
    for (long i = 0; i < n0; ++i) tmp[i] = (double)img_other[i] / 256.0;
    for (long i = 0; i < 2 * n0; ++i) tmp2[i] = (double)cvf[i];
    warping_image(others[0], tmp, 1, tmp2, H0, W0);
    for (int k = 1; k < L; ++k) {
        downsample_mask(masks[k], hs[k], ws[k], masks[k - 1], hs[k - 1], ws[k - 1]);
        downsample_bilinear(trs[k], hs[k], ws[k], trs[k - 1], hs[k - 1], ws[k - 1], 2);
        downsample_bilinear(refs[k], hs[k], ws[k], refs[k - 1], hs[k - 1], ws[k - 1], 1);
        downsample_bilinear(others[k], hs[k], ws[k], others[k - 1], hs[k - 1], ws[k - 1], 1);
    }
    for (int k = L; k >= 1; --k) {
        const long H = hs[k - 1], W = ws[k - 1], n = H * W;
        const uint8_t* msk = masks[k - 1];
        if (k == L) {
            memset(warp, 0, sizeof(double) * 2 * n);
            memset(u, 0, sizeof(double) * n);
            memset(ub, 0, sizeof(double) * n);
        }
        gaussian_masked(ref_s, refs[k - 1], msk, H, W);
        compute_tensor_eta(abc, eta, ref_s, msk, H, W, prm);
        for (int w = 0; w < NW; ++w) {
            warping_image(tv, trs[k - 1], 2, warp, H, W);
            warping_image(ow, others[k - 1], 1, warp, H, W);
            compute_derivatives(deriv, ow, refs[k - 1], trs[k - 1], H, W);
            memset(p, 0, sizeof(double) * 2 * n);
            memset(q, 0, sizeof(double) * 4 * n);
            memset(vb, 0, sizeof(double) * 2 * n);
            memcpy(u0, u, sizeof(double) * n);
            for (int it = 0; it < NI; ++it) {
                double* t;
                update_dual(p, tp, q, sigmas[it], ub, vb, abc, msk, H, W, prm, pnb, qnb);
                t = p, p = pnb, pnb = t;
                t = q, q = qnb, qnb = t;
                l1_thresholding(tmp, taus[it], tp, u, ub, deriv, eta, msk, H, W, prm->lambda);
                t = u, u = tmp, tmp = t;
                update_primal(taus[it], mus[it], p, q, u, ub, vb, abc, eta, msk, H, W, prm);
            }
            median_filter(uf, u, msk, H, W, prm->median_k);
            memcpy(u, u0, sizeof(double) * n);
            flow_with_clamp(warp, u, uf, tv, msk, n, prm->limit_range);
        }
        if (k > 1) {
            const long Hd = hs[k - 2], Wd = ws[k - 2];
            upsampling(tmp, Hd, Wd, u, H, W, 1, prm->f_scale);
            memcpy(ub, tmp, sizeof(double) * Hd * Wd);
            memcpy(u, tmp, sizeof(double) * Hd * Wd);
            upsampling(tmp, Hd, Wd, warp, H, W, 2, prm->f_scale);
            memcpy(warp, tmp, sizeof(double) * 2 * Hd * Wd);
        }
    }
    for (long i = 0; i < n0; ++i) {
        flow_out[2 * i] = masks[0][i] == 0 ? 0.0 : warp[2 * i];
        flow_out[2 * i + 1] = masks[0][i] == 0 ? 0.0 : warp[2 * i + 1];
    }
    rc = 0;
done:
    for (int k = 0; k < L; ++k) {
        free(masks[k]);
        free(trs[k]);
        free(refs[k]);
        free(others[k]);
    }
    free(warp);
    free(u);
    free(ub);
    free(ref_s);
    free(abc);
    free(eta);
    free(tv);
    free(ow);
    free(deriv);
    free(p);
    free(q);
    free(tp);
    free(vb);
    free(u0);
    free(uf);
    free(pnb);
    free(qnb);
    free(tmp);
    free(tmp2);
    return rc;
}
