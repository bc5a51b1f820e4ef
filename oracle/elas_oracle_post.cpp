/*
 * elas_oracle_post.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU definitions of the post-processing filters that the reference declares
 * but never implements: speckle removal (Elas.h:31-32, KernelsCommon.h:19-20),
 * gap interpolation (Elas.h:33, KernelsCommon.h:21), median and adaptive-mean
 * filters (Elas.h:34-35).  The reference pipeline ends at the left/right check
 * (Elas.cu:107-119), so these definitions are this project's own, written
 * from the ELAS paper (Geiger, Roser, Urtasun, ACCV 2010): there is no
 * reference output to match, and their parity is pinned only by invariants and
 * brute force (parity unpinned against the reference).
 *
 * All filters read a snapshot and write a new map (no in-place races), and
 * treat values < 0 as invalid.  Each is off unless its parameter enables it:
 *   speckle removal     speckle_size > 0
 *   gap interpolation   ipol_gap_width > 0
 *   adaptive mean       filter_adaptive_mean
 *   median              filter_median
 * and they run in that order on D1 (and on D2 unless postprocess_only_left).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "elas_hip.h"

extern "C" {

/* 4-connected segments of valid pixels, neighbours joined when
 * |d_a - d_b| <= sim; segments with fewer than max_size pixels become -1. */
void eo_speckle(float* D, int W, int H, float sim, int max_size) {
    const size_t n = (size_t)W * H;
    std::vector<int32_t> label(n, -1);
    std::vector<int32_t> stack, members;
    for (size_t s = 0; s < n; s++) {
        if (D[s] < 0 || label[s] >= 0) continue;
        stack.clear();
        members.clear();
        stack.push_back((int32_t)s);
        label[s] = (int32_t)s;
        while (!stack.empty()) {
            int32_t a = stack.back();
            stack.pop_back();
            members.push_back(a);
            int u = a % W, v = a / W;
            const int nu[4] = {u - 1, u + 1, u, u};
            const int nv[4] = {v, v, v - 1, v + 1};
            for (int k = 0; k < 4; k++) {
                if (nu[k] < 0 || nu[k] >= W || nv[k] < 0 || nv[k] >= H) continue;
                int32_t b = nv[k] * W + nu[k];
                if (label[b] >= 0 || D[b] < 0) continue;
                if (fabsf(D[a] - D[b]) <= sim) {
                    label[b] = (int32_t)s;
                    stack.push_back(b);
                }
            }
        }
        if ((int)members.size() < max_size)
            for (int32_t a : members) D[a] = -1;
    }
}

/* Fill runs of invalid pixels of length <= gap that have a valid pixel on
 * both sides: with the mean of the two when they differ by less than 3, else
 * with the smaller.  Rows first, then columns of the row-filled map. */
void eo_gap_interpolation(float* D, int W, int H, int gap) {
    for (int pass = 0; pass < 2; pass++) {
        const int lines = pass ? W : H, len = pass ? H : W;
        std::vector<float> src(D, D + (size_t)W * H);
        for (int l = 0; l < lines; l++) {
            auto at = [&](int i) -> size_t { return pass ? (size_t)i * W + l : (size_t)l * W + i; };
            int i = 0;
            while (i < len) {
                if (src[at(i)] >= 0) { i++; continue; }
                int j = i;
                while (j < len && src[at(j)] < 0) j++;
                /* invalid run [i, j) */
                if (i > 0 && j < len && j - i <= gap) {
                    float d1 = src[at(i - 1)], d2 = src[at(j)];
                    float f = fabsf(d1 - d2) < 3.0f ? (d1 + d2) * 0.5f : std::min(d1, d2);
                    for (int k = i; k < j; k++) D[at(k)] = f;
                }
                i = j;
            }
        }
    }
}

/* Adaptive mean: rows then columns, window of +-3 valid pixels, weight
 * max(0, 4 - |d_k - d_c|) (centre weight 4); invalid pixels stay invalid. */
void eo_adaptive_mean(float* D, int W, int H) {
    for (int pass = 0; pass < 2; pass++) {
        const int lines = pass ? W : H, len = pass ? H : W;
        std::vector<float> src(D, D + (size_t)W * H);
        for (int l = 0; l < lines; l++)
            for (int i = 0; i < len; i++) {
                size_t c = pass ? (size_t)i * W + l : (size_t)l * W + i;
                float dc = src[c];
                if (dc < 0) continue;
                float ws = 0, acc = 0;
                for (int k = -3; k <= 3; k++) {
                    int j = i + k;
                    if (j < 0 || j >= len) continue;
                    float dk = src[pass ? (size_t)j * W + l : (size_t)l * W + j];
                    if (dk < 0) continue;
                    float w = std::max(0.0f, 4.0f - fabsf(dk - dc));
                    ws += w;
                    acc += w * dk;
                }
                D[c] = acc / ws;
            }
    }
}

/* Median: rows then columns, window +-3, median of the valid values (the
 * lower middle element for even counts); invalid pixels stay invalid. */
void eo_median(float* D, int W, int H) {
    for (int pass = 0; pass < 2; pass++) {
        const int lines = pass ? W : H, len = pass ? H : W;
        std::vector<float> src(D, D + (size_t)W * H);
        float buf[7];
        for (int l = 0; l < lines; l++)
            for (int i = 0; i < len; i++) {
                size_t c = pass ? (size_t)i * W + l : (size_t)l * W + i;
                if (src[c] < 0) continue;
                int n = 0;
                for (int k = -3; k <= 3; k++) {
                    int j = i + k;
                    if (j < 0 || j >= len) continue;
                    float dk = src[pass ? (size_t)j * W + l : (size_t)l * W + j];
                    if (dk >= 0) buf[n++] = dk;
                }
                std::sort(buf, buf + n);
                D[c] = buf[(n - 1) / 2];
            }
    }
}

void eo_postprocess(float* D, int W, int H, const elas_params* p) {
    if (p->speckle_size > 0) eo_speckle(D, W, H, p->speckle_sim_threshold, p->speckle_size);
    if (p->ipol_gap_width > 0) eo_gap_interpolation(D, W, H, p->ipol_gap_width);
    if (p->filter_adaptive_mean) eo_adaptive_mean(D, W, H);
    if (p->filter_median) eo_median(D, W, H);
}

}  // extern "C"
