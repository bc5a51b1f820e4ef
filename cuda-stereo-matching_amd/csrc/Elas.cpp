// Elas.cpp -- the drop-in class Elas (include/elas/Elas.h) over the C ABI.
#include "elas/Elas.h"

#include <cstddef>
#include <cstring>

static_assert(sizeof(Elas::parameters) == sizeof(elas_params), "parameters layout");
static_assert(offsetof(Elas::parameters, add_corners) == offsetof(elas_params, add_corners), "layout");
static_assert(offsetof(Elas::parameters, grid_size) == offsetof(elas_params, grid_size), "layout");
static_assert(offsetof(Elas::parameters, subsampling) == offsetof(elas_params, subsampling), "layout");

Elas::parameters::parameters() {
    elas_params p;
    elas_params_default(&p);
    memcpy(static_cast<void*>(this), &p, sizeof(p));
}

Elas::parameters Elas::parameters::declared_defaults() {
    parameters r;
    elas_params p;
    elas_params_declared(&p);
    memcpy(static_cast<void*>(&r), &p, sizeof(p));
    return r;
}

Elas::Elas() : ctx_(nullptr), status_(ELAS_OK) {
    status_ = elas_hip_create(reinterpret_cast<const elas_params*>(&param), 0, &ctx_);
}

Elas::Elas(const parameters& prm, int device) : param(prm), ctx_(nullptr), status_(ELAS_OK) {
    status_ = elas_hip_create(reinterpret_cast<const elas_params*>(&param), device, &ctx_);
}

Elas::~Elas() {
    if (ctx_) elas_hip_destroy(ctx_);
}

void Elas::process(uint8_t* h_img0, uint8_t* h_img1, float* h_disp0, float* h_disp1, const int32_t* dims) {
    if (!ctx_) return;
    status_ = elas_hip_process(ctx_, h_img0, h_img1, h_disp0, h_disp1, dims);
}

const char* Elas::last_error() const { return ctx_ ? elas_hip_last_error(ctx_) : "no device context"; }
