/*
 * elas_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference ELAS path (beautifulv0id/CUDA-Stereo-Matching),
 * kernel by kernel, used as the parity checker for the HIP path and as the
 * CPU baseline of the benchmark.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository).  Where the reference has a race or undefined
 * behaviour, the convention encoded here is stated at the function:
 *   - support reduction: thread 0's own best is counted twice
 *     (KerSupportMatches.cu:154-178), restated literally as 32 simulated lanes;
 *   - ratio test in float: min1 < thr * (float)min2 (KerSupportMatches.cu:179);
 *   - inconsistency filter: Jacobi (reads a snapshot) instead of the racy
 *     in-place update of KerSupportMatches.cu:196-238;
 *   - descriptor bytes the reference never writes are 0 (KerDescriptor.cu:79);
 *   - overlapping triangle spans: the highest triangle index wins
 *     (KerComputeDisparity.cu:128-150 races); float->int casts saturate with
 *     NaN -> 0 (the GPU conversion), FP contraction is off (build flag);
 *   - every compile-time macro of KernelsCommon.h:8-27 is read from elas_params;
 *   - add_corners, declared but not implemented by the reference (Elas.h:23,
 *     "add support points at image corners with nearest neighbor
 *     disparities"), is this project's definition (eo_add_corners).
 *
 * Parity pins (tests/): closed forms for Sobel/descriptor/P/ratio test,
 * brute-force numpy restatements of the support rule (closed form of the
 * thread-0 quirk), of the filters, grid and dense argmin on small images,
 * invariants (shift images, constant images), the analytic plane, and the
 * empty-circle property of Triangle's meshes.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "elas_hip.h"

#define TRILIB elas_oracle_tri
#include "triangle.h"
#undef TRILIB

namespace {

inline int32_t iabs(int32_t x) { return x < 0 ? -x : x; }

/* float -> int32 with the GPU's saturating semantics (NaN -> 0). */
inline int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
/* float -> uint32 with saturation (NaN -> 0). */
inline uint32_t f2u_sat(float x) {
    if (x != x || x <= 0.0f) return 0u;
    if (x >= 4294967296.0f) return UINT32_MAX;
    return (uint32_t)x;
}

/* plane radius, KerComputeDisparity.cu:21 and :194 */
inline int32_t plane_radius(const elas_params* p) {
    return (int32_t)std::max((float)ceilf(p->sigma * p->sradius), 2.0f);
}

/* P(delta), KerComputeDisparity.cu:249-251,269-271 (computed per candidate
 * there; same expression and types here). */
inline int32_t prior_P(const elas_params* p, int32_t delta_d) {
    float two_sigma_squared = 2 * p->sigma * p->sigma;
    float e = expf((float)(-delta_d * delta_d) / two_sigma_squared);
    float num = -logf(p->gamma + e) + logf(p->gamma);
    return (int32_t)((double)num / (double)p->beta);
}

}  // namespace

extern "C" {

/* prior P(|d - d_plane|) and plane radius, exposed for tests */
int32_t eo_prior(const elas_params* p, int32_t delta) { return prior_P(p, delta); }
int32_t eo_plane_radius(const elas_params* p) { return plane_radius(p); }

/* ---- Sobel, KerDescriptor.cu:9-49 (kerComputeSobel) ---------------------- */
void eo_sobel(const uint8_t* img, int W, int H, uint8_t* du, uint8_t* dv) {
    for (int v = 1; v < H - 1; v++)
        for (int u = 1; u < W - 1; u++) {
            auto I = [&](int uu, int vv) { return (int32_t)img[vv * W + uu]; };
            float sum_du = (float)(I(u - 1, v - 1) + 2 * I(u - 1, v) + I(u - 1, v + 1)
                                   - I(u + 1, v - 1) - 2 * I(u + 1, v) - I(u + 1, v + 1));
            float sum_dv = (float)(I(u - 1, v - 1) + 2 * I(u, v - 1) + I(u + 1, v - 1)
                                   - I(u - 1, v + 1) - 2 * I(u, v + 1) - I(u + 1, v + 1));
            /* KerDescriptor.cu:40-41: (uint8_t)(min(max(-128, s/4), 127) + 128) */
            du[v * W + u] = (uint8_t)(std::min(std::max(-128.f, sum_du / 4), 127.f) + 128);
            dv[v * W + u] = (uint8_t)(std::min(std::max(-128.f, sum_dv / 4), 127.f) + 128);
        }
}

/* ---- Descriptor, KerDescriptor.cu:53-100 (kerComputeDescriptor) ---------- */
/* desc is [H][W][16]; pixels outside u in [3,W-4], v in [3,H-4] are zero. */
void eo_descriptor(const uint8_t* img, int W, int H, uint8_t* desc) {
    std::vector<uint8_t> du((size_t)W * H, 0), dv((size_t)W * H, 0);
    eo_sobel(img, W, H, du.data(), dv.data());
    memset(desc, 0, (size_t)16 * W * H);
    for (int v = 3; v < H - 3; v++)
        for (int u = 3; u < W - 3; u++) {
            uint8_t* o = desc + ((size_t)v * W + u) * 16;
            auto DU = [&](int uu, int vv) { return du[(size_t)vv * W + uu]; };
            auto DV = [&](int uu, int vv) { return dv[(size_t)vv * W + uu]; };
            o[0] = DU(u, v - 2);
            o[1] = DU(u - 2, v - 1); o[2] = DU(u, v - 1); o[3] = DU(u + 2, v - 1);
            o[4] = DU(u - 1, v); o[5] = DU(u, v); o[6] = DU(u, v); o[7] = DU(u + 1, v);
            o[8] = DU(u - 2, v + 1); o[9] = DU(u, v + 1); o[10] = DU(u + 2, v + 1);
            o[11] = DU(u, v + 2);
            o[12] = DV(u, v - 1); o[13] = DV(u - 1, v); o[14] = DV(u + 1, v); o[15] = DV(u, v + 1);
        }
}

/* ---- Support matching, KerSupportMatches.cu:13-192 ----------------------- */
/* One pass over the candidate lattice.  Left pass (right=0) writes every
 * in-bounds candidate; cells outside the bounds keep their value (the
 * caller initialises can to -1, Elas.cu:190).  Right pass (right=1) reads
 * the left result and invalidates cells failing the LR test.  *work counts
 * the (candidate, d) pairs whose 64-byte SAD is evaluated. */
void eo_support_pass(const uint8_t* desc0, const uint8_t* desc1, int W, int H,
                     const elas_params* p, int right, int32_t* can, int64_t* work) {
    const int step = p->candidate_stepsize;
    const int cw = W / step, ch = H / step;
    const int x_step = 2, y_step = 2, desc_window = 3;
    const int tpc = 32;
    const long long off[4] = {-16LL * x_step - 16LL * W * y_step, +16LL * x_step - 16LL * W * y_step,
                              -16LL * x_step + 16LL * W * y_step, +16LL * x_step + 16LL * W * y_step};
    for (int v_can = 0; v_can < ch; v_can++)
        for (int u_can = 0; u_can < cw; u_can++) {
            int32_t u = u_can * step, v = v_can * step;
            int32_t* cell = &can[v_can * cw + u_can];
            if (!(u >= x_step + desc_window && u <= W - x_step - desc_window - 1 &&
                  v >= y_step + desc_window && v <= H - y_step - desc_window - 1))
                continue;
            int32_t d_left = 0;
            if (right) {
                d_left = *cell;
                if (d_left < 0) continue;
                u = u - d_left;
            }
            const uint8_t* I1 = (right ? desc1 : desc0) + (size_t)16 * W * v;
            const uint8_t* I2 = (right ? desc0 : desc1) + (size_t)16 * W * v;
            const uint8_t* blk = I1 + 16 * (long long)u;
            int32_t sum = 0;
            for (int i = 0; i < 16; i++) sum += iabs((int32_t)blk[i] - 128);
            if (sum < p->support_texture) { *cell = -1; continue; }
            int32_t disp_min_valid = std::max(p->disp_min, 0);
            int32_t disp_max_valid = !right ? std::min(p->disp_max, u - desc_window - x_step)
                                            : std::min(p->disp_max, W - u - desc_window - x_step);
            if (disp_max_valid - disp_min_valid < 10) { *cell = -1; continue; }
            if (work) *work += disp_max_valid - disp_min_valid + 1;
            /* 32 simulated threads, each strided over d (:124-152) */
            int32_t sh[4 * 32];
            int16_t t0_m1E = 32767, t0_m1d = -1, t0_m2E = 32767, t0_m2d = -1;
            for (int t = 0; t < tpc; t++) {
                int16_t m1E = 32767, m1d = -1, m2E = 32767, m2d = -1;
                for (int32_t d = disp_min_valid + t; d <= disp_max_valid; d += tpc) {
                    int32_t u_warp = !right ? u - d : u + d;
                    const uint8_t* b2 = I2 + 16 * (long long)u_warp;
                    int32_t s = 0;
                    for (int i = 0; i < 4; i++)
                        for (int j = 0; j < 16; j++)
                            s += iabs((int32_t)blk[off[i] + j] - (int32_t)b2[off[i] + j]);
                    if (s < m1E) { m2E = m1E; m2d = m1d; m1E = (int16_t)s; m1d = (int16_t)d; }
                    else if (s < m2E) { m2E = (int16_t)s; m2d = (int16_t)d; }
                }
                sh[4 * t + 0] = m1E; sh[4 * t + 1] = m1d; sh[4 * t + 2] = m2E; sh[4 * t + 3] = m2d;
                if (t == 0) { t0_m1E = m1E; t0_m1d = m1d; t0_m2E = m2E; t0_m2d = m2d; }
            }
            /* thread 0 reduction starting from its own registers (:161-175) */
            int32_t min_1_E = t0_m1E, min_1_d = t0_m1d, min_2_E = t0_m2E, min_2_d = t0_m2d;
            for (int i = 0; i < 4 * tpc; i += 2) {
                int32_t tE = sh[i], td = sh[i + 1];
                if (tE < min_1_E) { min_2_E = min_1_E; min_2_d = min_1_d; min_1_E = tE; min_1_d = td; }
                else if (tE < min_2_E) { min_2_E = tE; min_2_d = td; }
            }
            /* ratio test in float (:179); LR test on the right pass (:180-188) */
            float thr_m2 = p->support_threshold * (float)min_2_E;
            if (min_1_d >= 0 && min_2_d >= 0 && (float)min_1_E < thr_m2) {
                if (!right) *cell = min_1_d;
                else if (iabs(d_left - min_1_d) > p->lr_threshold) *cell = -1;
            } else {
                *cell = -1;
            }
        }
}

/* ---- Inconsistency filter, KerSupportMatches.cu:196-238 ------------------ */
/* Jacobi: reads `in`, writes `out`.  Row guard y_can < can_width as at :208. */
void eo_incon_filter(const int32_t* in, int32_t* out, int cw, int ch, const elas_params* p) {
    memcpy(out, in, sizeof(int32_t) * (size_t)cw * ch);
    const int R = p->incon_window_size;
    for (int y = 0; y < ch; y++)
        for (int x = 0; x < cw; x++) {
            if (!(x < cw && y < cw)) continue;
            int32_t d = in[y * cw + x];
            if (d < 0) continue;
            int32_t support = 0;
            for (int x2 = x - R; x2 <= x + R; x2++)
                for (int y2 = y - R; y2 <= y + R; y2++)
                    if (x2 >= 0 && x2 < cw && y2 >= 0 && y2 < ch) {
                        int32_t d2 = in[y2 * cw + x2];
                        if (d2 >= 0 && iabs(d - d2) <= p->incon_threshold) support++;
                    }
            if (support < p->incon_min_support) out[y * cw + x] = -1;
        }
}

/* ---- Redundancy filters, KerSupportMatches.cu:242-356 -------------------- */
/* Sequential in-place scans: every row left to right, then every column top
 * to bottom (Elas.cu:248-272).  REDUN_MAX_DIST 5, REDUN_THRESHOLD 1. */
void eo_redun_filter(int32_t* can, int cw, int ch) {
    const int max_dist = 5, thr = 1;
    for (int v = 0; v < ch; v++)
        for (int u = 0; u < cw; u++) {
            int32_t d = can[v * cw + u];
            if (d < 0) continue;
            bool redundant = true;
            for (int dir = -1; dir <= 1; dir += 2) {
                int u2 = u;
                bool support = false;
                for (int j = 0; j < max_dist; j++) {
                    u2 += dir;
                    if (u2 < 0 || u2 >= cw) break;
                    int32_t d2 = can[v * cw + u2];
                    if (d2 >= 0 && iabs(d - d2) <= thr) { support = true; break; }
                }
                if (!support) { redundant = false; break; }
            }
            if (redundant) can[v * cw + u] = -1;
        }
    for (int u = 0; u < cw; u++)
        for (int v = 0; v < ch; v++) {
            int32_t d = can[v * cw + u];
            if (d < 0) continue;
            bool redundant = true;
            for (int dir = -1; dir <= 1; dir += 2) {
                int v2 = v;
                bool support = false;
                for (int j = 0; j < max_dist; j++) {
                    v2 += dir;
                    if (v2 < 0 || v2 >= ch) break;
                    int32_t d2 = can[v2 * cw + u];
                    if (d2 >= 0 && iabs(d - d2) <= thr) { support = true; break; }
                }
                if (!support) { redundant = false; break; }
            }
            if (redundant) can[v * cw + u] = -1;
        }
}

/* ---- Compaction, KerSupportMatches.cu:361-423 + Elas.cu:338-394 ---------- */
/* Row-major order; sp0 = (u*step, v*step), sp1 = (u*step - d, v*step). */
int eo_compact(const int32_t* can, int cw, int ch, int step, float* sp0, float* sp1) {
    int n = 0;
    for (int v = 0; v < ch; v++)
        for (int u = 0; u < cw; u++) {
            int32_t d = can[v * cw + u];
            if (d < 0) continue;
            float pu = (float)(u * step), pv = (float)(v * step);
            if (sp0) { sp0[2 * n] = pu; sp0[2 * n + 1] = pv; }
            if (sp1) { sp1[2 * n] = pu - d; sp1[2 * n + 1] = pv; }
            n++;
        }
    return n;
}

/* ---- Corner support points (add_corners, Elas.h:23) ---------------------------
 * The reference declares the switch and never implements it; its field
 * comment is the specification.  Definition used here and on the GPU: each
 * image corner (0, 0), (0, H-1), (W-1, 0), (W-1, H-1) becomes a support point
 * with the disparity d of the nearest support point (smallest squared
 * distance in pixels to its left-image position, the first in compaction
 * order on ties), appended after the nsp support points in that order; then,
 * so that the right image's mesh also reaches its right border, the two right
 * corners again at u = W-1+d (their right-image position is u - d = W-1).
 * Each point is (u, v) in sp0 and (u - d, v) in sp1.  Only when nsp >= 1;
 * returns the new count. */
int eo_add_corners(float* sp0, float* sp1, int nsp, int W, int H) {
    if (nsp < 1) return nsp;
    const int cu[4] = {0, 0, W - 1, W - 1}, cv[4] = {0, H - 1, 0, H - 1};
    int cd[4];
    for (int c = 0; c < 4; c++) {
        long best = -1;
        for (int i = 0; i < nsp; i++) {
            const long du = (long)sp0[2 * i] - cu[c], dv = (long)sp0[2 * i + 1] - cv[c];
            const long dist = du * du + dv * dv;
            if (best < 0 || dist < best) {
                best = dist;
                cd[c] = (int)(sp0[2 * i] - sp1[2 * i]);
            }
        }
    }
    const int u[6] = {cu[0], cu[1], cu[2], cu[3], W - 1 + cd[2], W - 1 + cd[3]};
    const int v[6] = {cv[0], cv[1], cv[2], cv[3], 0, H - 1};
    const int d[6] = {cd[0], cd[1], cd[2], cd[3], cd[2], cd[3]};
    for (int k = 0; k < 6; k++) {
        sp0[2 * (nsp + k)] = (float)u[k];
        sp0[2 * (nsp + k) + 1] = (float)v[k];
        sp1[2 * (nsp + k)] = (float)(u[k] - d[k]);
        sp1[2 * (nsp + k) + 1] = (float)v[k];
    }
    return nsp + 6;
}

/* ---- Disparity grid, KerGrid.cu:9-126 + Elas.cu:403-464 ------------------ */
void eo_grid_dims(int W, int H, const elas_params* p, int* gw, int* gh) {
    *gw = (int32_t)ceil((float)W / (float)p->grid_size);
    *gh = (int32_t)ceil((float)H / (float)p->grid_size);
}

void eo_grid(const int32_t* can, int cw, int ch, int W, int H, const elas_params* p,
             int32_t* grid0, int32_t* grid1) {
    int gw, gh;
    eo_grid_dims(W, H, p, &gw, &gh);
    const int D = p->disp_max + 1;
    const int step = p->candidate_stepsize;
    const float gs = (float)p->grid_size;
    std::vector<uint8_t> t0((size_t)gw * gh * D, 0), t1((size_t)gw * gh * D, 0);
    /* kerLoadGridValues :9-53 */
    for (int v_can = 0; v_can < ch; v_can++)
        for (int u_can = 0; u_can < cw; u_can++) {
            int32_t d_curr = can[v_can * cw + u_can];
            if (d_curr < 0) continue;
            int32_t d_min = std::max(d_curr - 1, 0), d_max = std::min(d_curr + 1, p->disp_max);
            int32_t u_img = u_can * step, v_img = v_can * step;
            for (int d = d_min; d <= d_max; d++) {
                int32_t u_grid0 = (int32_t)floorf((float)u_img / gs);
                int32_t u_grid1 = (int32_t)floorf((float)(u_img - d_curr) / gs);
                int32_t v_grid = (int32_t)floorf((float)v_img / gs);
                if (v_grid >= 0 && v_grid < gh) {
                    if (u_grid0 >= 0 && u_grid0 < gw) t0[((size_t)v_grid * gw + u_grid0) * D + d] = 1;
                    if (u_grid1 >= 0 && u_grid1 < gw) t1[((size_t)v_grid * gw + u_grid1) * D + d] = 1;
                }
            }
        }
    /* kerCreateGird :57-126: interior cells only; 3x3 OR then ascending list */
    const size_t cell = (size_t)D + 1;
    for (int k = 0; k < 2; k++) {
        const std::vector<uint8_t>& t = k ? t1 : t0;
        int32_t* g = k ? grid1 : grid0;
        memset(g, 0, sizeof(int32_t) * (size_t)gw * gh * cell);
        for (int v = 1; v < gh - 1; v++)
            for (int u = 1; u < gw - 1; u++) {
                int32_t* out = g + ((size_t)v * gw + u) * cell;
                int32_t n = 0;
                for (int d = 0; d < D; d++) {
                    uint8_t f = 0;
                    for (int dv = -1; dv <= 1; dv++)
                        for (int du = -1; du <= 1; du++)
                            f |= t[((size_t)(v + dv) * gw + (u + du)) * D + d];
                    if (f > 0) out[1 + n++] = d;
                }
                out[0] = n;
            }
    }
}

/* ---- Delaunay: TRILIB::triangulate("zQB"), Elas.cu:282-328 --------------- */
/* Returns the triangle count, or -1 with fewer than 3 points (the reference
 * exits) or when cap is too small. */
int eo_triangulate(const float* sp, int n, int32_t* tri, int cap) {
    if (n < 3) return -1;
    elas_oracle_tri::triangulateio in, out;
    memset(&in, 0, sizeof(in));
    memset(&out, 0, sizeof(out));
    float* pts = (float*)malloc(sizeof(float) * 2 * (size_t)n);
    memcpy(pts, sp, sizeof(float) * 2 * (size_t)n);
    in.numberofpoints = n;
    in.pointlist = pts;
    char sw[] = "zQB";
    elas_oracle_tri::triangulate(sw, &in, &out, nullptr);
    int nt = out.numberoftriangles;
    int ret = nt;
    if (nt > cap) ret = -1;
    else if (tri) memcpy(tri, out.trianglelist, sizeof(int32_t) * 3 * (size_t)nt);
    free(pts);
    free(out.pointlist);
    free(out.trianglelist);
    return ret;
}

/* ---- Plane prior, KerComputeDisparity.cu:10-152 + kerSolve :322-375 ------ */
static bool solve3(double A[3][3], double B[3]) {
    const double eps = 1e-20;
    int ipiv[3] = {0, 0, 0};
    int irow = 0, icol = 0;
    for (int i = 0; i < 3; i++) {
        double big = 0.0;
        for (int j = 0; j < 3; j++)
            if (ipiv[j] != 1)
                for (int k = 0; k < 3; k++)
                    if (ipiv[k] == 0)
                        if (fabs(A[j][k]) >= big) { big = fabs(A[j][k]); irow = j; icol = k; }
        ++(ipiv[icol]);
        if (irow != icol) {
            for (int l = 0; l < 3; l++) std::swap(A[irow][l], A[icol][l]);
            std::swap(B[irow], B[icol]);
        }
        if (fabs(A[icol][icol]) < eps) return false;
        double pivinv = 1.0 / A[icol][icol];
        A[icol][icol] = 1.0;
        for (int l = 0; l < 3; l++) A[icol][l] *= pivinv;
        B[icol] *= pivinv;
        for (int ll = 0; ll < 3; ll++)
            if (ll != icol) {
                double dum = A[ll][icol];
                A[ll][icol] = 0.0;
                for (int l = 0; l < 3; l++) A[ll][l] -= A[icol][l] * dum;
                B[ll] -= B[icol] * dum;
            }
    }
    return true;
}

/* A support point's disparity: from the lattice at its position
 * (KerComputeDisparity.cu:54-56), or, when sp1 is given, as sp0.u - sp1.u,
 * which is the same number for lattice points and the only one for the
 * corner points of eo_add_corners. */
static inline int32_t point_d(const int32_t* can, int cw, int step, const float* sp, const float* sp1, int j) {
    if (sp1) return (int32_t)(sp[2 * j] - sp1[2 * j]);
    return can[((int32_t)sp[2 * j + 1] * cw + (int32_t)sp[2 * j]) / step];
}

static void plane_of(const int32_t* can, int cw, int step, const float* sp, const float* sp1, const int32_t* t,
                     int right, float* abc) {
    int32_t tu[3], tv[3], td[3];
    double A[3][3], b[3];
    for (int i = 0; i < 3; i++) {
        tu[i] = (int32_t)sp[2 * t[i]];
        tv[i] = (int32_t)sp[2 * t[i] + 1];
        td[i] = point_d(can, cw, step, sp, sp1, t[i]);
        A[i][0] = right ? (double)(tu[i] - td[i]) : (double)tu[i];
        A[i][1] = tv[i];
        A[i][2] = 1;
        b[i] = td[i];
    }
    if (solve3(A, b)) { abc[0] = (float)b[0]; abc[1] = (float)b[1]; abc[2] = (float)b[2]; }
    else { abc[0] = 0; abc[1] = 0; abc[2] = 0; }
}

/* Plane (a,b,c) of one triangle as floats; right=1 shifts corners by -d. */
void eo_plane(const int32_t* can, int cw, int step, const float* sp, const int32_t* t, int right,
              float* abc) {
    plane_of(can, cw, step, sp, nullptr, t, right, abc);
}

/* Rasterise every triangle in index order into mean (W*H, caller-initialised
 * to -1): a later triangle overwrites an earlier one (highest index wins).
 * Span rows are clamped to [0, H) (the reference would write out of bounds).
 * With sp1, point disparities are sp0.u - sp1.u (see point_d). */
void eo_raster_sp(const int32_t* can, int cw, const float* sp0, const float* sp1, const int32_t* tri, int ntri,
                  int right, int W, int H, const elas_params* p, float* mean) {
    const int step = p->candidate_stepsize;
    for (int t = 0; t < ntri; t++) {
        float abc[3];
        plane_of(can, cw, step, sp0, sp1, tri + 3 * t, right, abc);
        const float pa = abc[0], pb = abc[1], pc = abc[2];
        int32_t tu[3], tv[3], td[3];
        for (int i = 0; i < 3; i++) {
            tu[i] = (int32_t)sp0[2 * tri[3 * t + i]];
            tv[i] = (int32_t)sp0[2 * tri[3 * t + i] + 1];
            td[i] = point_d(can, cw, step, sp0, sp1, tri[3 * t + i]);
            if (right) tu[i] -= td[i];
        }
        /* insertion-style sort by u, :101-110 */
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < j; k++)
                if (tu[k] > tu[j]) { std::swap(tu[j], tu[k]); std::swap(tv[j], tv[k]); }
        float A_u = (float)tu[0], A_v = (float)tv[0];
        float B_u = (float)tu[1], B_v = (float)tv[1];
        float C_u = (float)tu[2], C_v = (float)tv[2];
        float AB_a = 0, AC_a = 0, BC_a = 0;
        if ((int32_t)A_u != (int32_t)B_u) AB_a = (A_v - B_v) / (A_u - B_u);
        if ((int32_t)A_u != (int32_t)C_u) AC_a = (A_v - C_v) / (A_u - C_u);
        if ((int32_t)B_u != (int32_t)C_u) BC_a = (B_v - C_v) / (B_u - C_u);
        float AB_b = A_v - AB_a * A_u;
        float AC_b = A_v - AC_a * A_u;
        float BC_b = B_v - BC_a * B_u;
        for (int part = 0; part < 2; part++) {
            float lo_u = part ? B_u : A_u, hi_u = part ? C_u : B_u;
            if ((int32_t)lo_u == (int32_t)hi_u) continue;
            float ea = part ? BC_a : AB_a, eb = part ? BC_b : AB_b;
            for (int32_t u = std::max((int32_t)lo_u, 0); u < std::min((int32_t)hi_u, W); u++) {
                int32_t v_1 = (int32_t)f2u_sat(AC_a * (float)u + AC_b);
                int32_t v_2 = (int32_t)f2u_sat(ea * (float)u + eb);
                int32_t v0 = std::max(std::min(v_1, v_2), 0), v1 = std::min(std::max(v_1, v_2), H);
                for (int32_t v = v0; v < v1; v++) {
                    int32_t d_plane = f2i_sat(pa * (float)u + pb * (float)v + pc);
                    mean[(size_t)v * W + u] = (float)d_plane;
                }
            }
        }
    }
}

void eo_raster(const int32_t* can, int cw, const float* sp0, const int32_t* tri, int ntri, int right,
               int W, int H, const elas_params* p, float* mean) {
    eo_raster_sp(can, cw, sp0, nullptr, tri, ntri, right, W, H, p, mean);
}

/* ---- Dense matching, KerComputeDisparity.cu:156-316,378-402 -------------- */
/* disp (W*H) is fully written: -1 where no disparity is computed.  *work
 * counts the (pixel, d) pairs whose 16-byte SAD is evaluated. */
void eo_dense(const int32_t* grid, const uint8_t* desc0, const uint8_t* desc1, const float* mean,
              int right, int W, int H, const elas_params* p, float* disp, int64_t* work) {
    int gw, gh;
    eo_grid_dims(W, H, p, &gw, &gh);
    const int desc_window = 3;
    const int disp_num = p->disp_max + 1;
    const int pr = plane_radius(p);
    for (size_t i = 0; i < (size_t)W * H; i++) disp[i] = -1;
    for (int v = desc_window + 1; v < H - desc_window; v++) {
        const uint8_t* I1 = (right ? desc1 : desc0) + (size_t)16 * W * v;
        const uint8_t* I2 = (right ? desc0 : desc1) + (size_t)16 * W * v;
        for (int u = desc_window + 1; u < W - desc_window; u++) {
            int32_t d_plane = f2i_sat(mean[(size_t)v * W + u]);
            if (d_plane <= 0) continue;
            const uint8_t* b1 = I1 + 16 * (size_t)u;
            int32_t sum = 0;
            for (int i = 0; i < 16; i++) sum += iabs((int32_t)b1[i] - 128);
            if (sum < p->match_texture) continue; /* the reference leaves the whole thread (:231-232) */
            int32_t dpmin = std::max(d_plane - pr, 0), dpmax = std::min(d_plane + pr, disp_num - 1);
            int32_t gx = (int32_t)floorf((float)u / (float)p->grid_size);
            int32_t gy = (int32_t)floorf((float)v / (float)p->grid_size);
            const int32_t* g = grid + ((size_t)gy * gw + gx) * (size_t)(p->disp_max + 2);
            int32_t num_grid = g[0];
            int32_t min_val = 10000, min_d = -1;
            auto sad = [&](int32_t u_warp) {
                const uint8_t* b2 = I2 + 16 * (size_t)u_warp;
                int32_t s = 0;
                for (int j = 0; j < 16; j++) s += iabs((int32_t)b2[j] - (int32_t)b1[j]);
                return s;
            };
            for (int i = 0; i < num_grid; i++) {
                int32_t d_curr = g[1 + i];
                if (d_curr < dpmin || d_curr > dpmax) {
                    int32_t u_warp = right ? u + d_curr : u - d_curr;
                    if (u_warp < desc_window || u_warp >= W - desc_window) continue;
                    if (work) (*work)++;
                    int32_t val = sad(u_warp);
                    if (val < min_val) { min_val = val; min_d = d_curr; }
                }
            }
            for (int32_t d_curr = dpmin; d_curr <= dpmax; d_curr++) {
                int32_t u_warp = right ? u + d_curr : u - d_curr;
                if (u_warp < desc_window || u_warp >= W - desc_window) continue;
                if (work) (*work)++;
                int32_t val = prior_P(p, iabs(d_curr - d_plane)) + sad(u_warp);
                if (val < min_val) { min_val = val; min_d = d_curr; }
            }
            disp[(size_t)v * W + u] = min_d >= 0 ? (float)min_d : -1.0f;
        }
    }
}

/* ---- Left/right check, KerLeftRightConsistencyCheck.cu:8-66 -------------- */
/* Reads copies of D1/D2 (Elas.cu:541-558), writes D1/D2 in place. */
void eo_lr(float* D1, float* D2, int W, int H, const elas_params* p) {
    std::vector<float> c1(D1, D1 + (size_t)W * H), c2(D2, D2 + (size_t)W * H);
    const float thr = (float)p->lr_threshold;
    for (int v = 0; v < H; v++)
        for (int u = 0; u < W; u++) {
            size_t a = (size_t)v * W + u;
            float d1 = c1[a], d2 = c2[a];
            float uw1 = (float)u - d1, uw2 = (float)u + d2;
            if (d1 >= 0 && uw1 >= 0 && uw1 < W) {
                if (fabsf(c2[(size_t)v * W + (int32_t)uw1] - d1) > thr) D1[a] = -1;
            } else {
                D1[a] = -1;
            }
            if (d2 >= 0 && uw2 >= 0 && uw2 < W) {
                if (fabsf(c1[(size_t)v * W + (int32_t)uw2] - d2) > thr) D2[a] = -1;
            } else {
                D2[a] = -1;
            }
        }
}

/* ---- Whole pipeline, Elas.cu:19-120 -------------------------------------- */
typedef struct eo_dump {
    uint8_t* desc0; uint8_t* desc1;                 /* 16*W*H each                 */
    int32_t* can_left; int32_t* can_lr;             /* cw*ch each                  */
    int32_t* can_incon; int32_t* can_final;
    float* sp0; float* sp1;                         /* 2*cw*ch capacity each       */
    int32_t* grid0; int32_t* grid1;                 /* gw*gh*(disp_max+2) each     */
    int32_t* tri0; int32_t* tri1; int32_t tri_cap;  /* 3*tri_cap each              */
    float* mean0; float* mean1;                     /* W*H each                    */
    float* d1_raw; float* d2_raw;                   /* W*H each                    */
    int32_t nsp, ntri0, ntri1;                      /* out                         */
    int64_t work_support, work_dense;               /* out                         */
} eo_dump;

void eo_postprocess(float* D, int W, int H, const elas_params* p);

/* Returns 0, or ELAS_ERR_TOO_FEW_POINTS (D1/D2 all -1) when Triangle would
 * exit, or ELAS_ERR_INVALID_ARGUMENT. */
int eo_process(const uint8_t* L, const uint8_t* R, int W, int H, const elas_params* p, float* D1,
               float* D2, eo_dump* dump) {
    if (W < 1 || H < 1 || !L || !R || !D1 || !D2 || p->candidate_stepsize < 1 || p->grid_size < 1)
        return ELAS_ERR_INVALID_ARGUMENT;
    const int step = p->candidate_stepsize;
    const int cw = W / step, ch = H / step;
    int gw, gh;
    eo_grid_dims(W, H, p, &gw, &gh);
    const size_t npx = (size_t)W * H;
    std::vector<uint8_t> desc0(16 * npx), desc1(16 * npx);
    eo_descriptor(L, W, H, desc0.data());
    eo_descriptor(R, W, H, desc1.data());
    std::vector<int32_t> can((size_t)cw * ch, -1), can2((size_t)cw * ch);
    int64_t wsup = 0, wden = 0;
    eo_support_pass(desc0.data(), desc1.data(), W, H, p, 0, can.data(), &wsup);
    if (dump && dump->can_left) memcpy(dump->can_left, can.data(), 4 * can.size());
    eo_support_pass(desc0.data(), desc1.data(), W, H, p, 1, can.data(), &wsup);
    if (dump && dump->can_lr) memcpy(dump->can_lr, can.data(), 4 * can.size());
    eo_incon_filter(can.data(), can2.data(), cw, ch, p);
    if (dump && dump->can_incon) memcpy(dump->can_incon, can2.data(), 4 * can.size());
    eo_redun_filter(can2.data(), cw, ch);
    if (dump && dump->can_final) memcpy(dump->can_final, can2.data(), 4 * can.size());
    std::vector<float> sp0(2 * (size_t)cw * ch + 16), sp1(2 * (size_t)cw * ch + 16);
    int nsp = eo_compact(can2.data(), cw, ch, step, sp0.data(), sp1.data());
    if (p->add_corners) nsp = eo_add_corners(sp0.data(), sp1.data(), nsp, W, H);
    const size_t gcell = (size_t)p->disp_max + 2;
    std::vector<int32_t> grid0((size_t)gw * gh * gcell), grid1((size_t)gw * gh * gcell);
    eo_grid(can2.data(), cw, ch, W, H, p, grid0.data(), grid1.data());
    if (dump) {
        if (dump->desc0) memcpy(dump->desc0, desc0.data(), desc0.size());
        if (dump->desc1) memcpy(dump->desc1, desc1.data(), desc1.size());
        if (dump->sp0) memcpy(dump->sp0, sp0.data(), sizeof(float) * 2 * nsp);
        if (dump->sp1) memcpy(dump->sp1, sp1.data(), sizeof(float) * 2 * nsp);
        if (dump->grid0) memcpy(dump->grid0, grid0.data(), 4 * grid0.size());
        if (dump->grid1) memcpy(dump->grid1, grid1.data(), 4 * grid1.size());
        dump->nsp = nsp;
    }
    for (size_t i = 0; i < npx; i++) { D1[i] = -1; D2[i] = -1; }
    if (nsp < 3) {
        if (dump) { dump->ntri0 = dump->ntri1 = 0; dump->work_support = wsup; dump->work_dense = 0; }
        return ELAS_ERR_TOO_FEW_POINTS;
    }
    const int cap = 2 * nsp + 8;
    std::vector<int32_t> tri0(3 * (size_t)cap), tri1(3 * (size_t)cap);
    int nt0 = eo_triangulate(sp0.data(), nsp, tri0.data(), cap);
    int nt1 = eo_triangulate(sp1.data(), nsp, tri1.data(), cap);
    if (nt0 < 0 || nt1 < 0) return ELAS_ERR_TOO_FEW_POINTS;
    std::vector<float> mean0(npx, -1.0f), mean1(npx, -1.0f);
    const float* dsrc = p->add_corners ? sp1.data() : nullptr;  // corners carry their own d
    eo_raster_sp(can2.data(), cw, sp0.data(), dsrc, tri0.data(), nt0, 0, W, H, p, mean0.data());
    eo_raster_sp(can2.data(), cw, sp0.data(), dsrc, tri1.data(), nt1, 1, W, H, p, mean1.data());
    eo_dense(grid0.data(), desc0.data(), desc1.data(), mean0.data(), 0, W, H, p, D1, &wden);
    eo_dense(grid1.data(), desc0.data(), desc1.data(), mean1.data(), 1, W, H, p, D2, &wden);
    if (dump) {
        if (dump->tri0 && nt0 <= dump->tri_cap) memcpy(dump->tri0, tri0.data(), 12 * (size_t)nt0);
        if (dump->tri1 && nt1 <= dump->tri_cap) memcpy(dump->tri1, tri1.data(), 12 * (size_t)nt1);
        if (dump->mean0) memcpy(dump->mean0, mean0.data(), 4 * npx);
        if (dump->mean1) memcpy(dump->mean1, mean1.data(), 4 * npx);
        if (dump->d1_raw) memcpy(dump->d1_raw, D1, 4 * npx);
        if (dump->d2_raw) memcpy(dump->d2_raw, D2, 4 * npx);
        dump->ntri0 = nt0;
        dump->ntri1 = nt1;
        dump->work_support = wsup;
        dump->work_dense = wden;
    }
    eo_lr(D1, D2, W, H, p);
    eo_postprocess(D1, W, H, p);
    if (!p->postprocess_only_left) eo_postprocess(D2, W, H, p);
    return 0;
}

}  // extern "C"
