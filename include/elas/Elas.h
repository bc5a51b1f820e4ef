/*
 * Elas.h -- drop-in replacement for the reference's class Elas
 * (src/Elas.h:11-167), running on one gfx950 GPU through the C ABI of
 * elas_hip.h.
 *
 *   Elas::parameters  same fields, types and order as src/Elas.h:14-67; the
 *                     default constructor gives the values the reference
 *                     kernels actually use (src/Kernels/KernelsCommon.h:8-27)
 *                     with post-processing off; declared_defaults() gives the
 *                     initialisers written at src/Elas.h:42-66.
 *   Elas()            device 0, effective defaults (src/Elas.h:70-72).
 *   Elas(parameters)  added: the reference's param member has no setter.
 *   process(...)      src/Elas.h:76-83.  Returns void like the reference; the
 *                     status of the last call is status(), and an error never
 *                     exits the process (the reference calls exit()).
 */
#ifndef ELAS_HIP_ELAS_H
#define ELAS_HIP_ELAS_H

#include <stdint.h>

#include "elas_hip.h"

class Elas {
public:
    struct parameters {
        int32_t disp_min;
        int32_t disp_max;
        float   support_threshold;
        int32_t support_texture;
        int32_t candidate_stepsize;
        int32_t incon_window_size;
        int32_t incon_threshold;
        int32_t incon_min_support;
        bool    add_corners;
        int32_t grid_size;
        float   beta;
        float   gamma;
        float   sigma;
        float   sradius;
        int32_t match_texture;
        int32_t lr_threshold;
        float   speckle_sim_threshold;
        int32_t speckle_size;
        int32_t ipol_gap_width;
        bool    filter_median;
        bool    filter_adaptive_mean;
        bool    postprocess_only_left;
        bool    subsampling;

        parameters();
        static parameters declared_defaults();
    };

    Elas();
    explicit Elas(const parameters& param, int device = 0);
    ~Elas();
    Elas(const Elas&) = delete;
    Elas& operator=(const Elas&) = delete;

    void process(uint8_t* h_img0, uint8_t* h_img1, float* h_disp0, float* h_disp1, const int32_t* dims);

    int status() const { return status_; }
    const char* last_error() const;
    elas_ctx* context() { return ctx_; }

private:
    parameters param;
    elas_ctx* ctx_;
    int status_;
};

#endif
