// r3d_gpu.hpp - header-only C++ layer over the C ABI of r3d_hip.h.
//
// Mirrors the reference's device-resident driver surface
// (reference include/gpu.hpp:35-49) for C++ callers: one Device object per
// run and GPU owns the context (gpu_alloc / gpu_free), moves the reference's
// per-rank arrays in and out (gpu_upload / gpu_download) and advances the
// state by whole or truncated RK3 steps (both step_gpu overloads).  Errors
// throw r3d_amd::Error carrying the R3D_E* code and the library's message;
// a caller that wants the reference's behaviour (message + exit) catches it
// at the top level.
#ifndef R3D_GPU_HPP
#define R3D_GPU_HPP

#include <stdexcept>
#include <string>

#include "r3d_hip.h"

namespace r3d_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

inline void check(int rc) {
    if (rc != R3D_OK) throw Error(rc, r3d_hip_last_error());
}

class Device {
   public:
    // gpu_alloc: validates the configuration and allocates X (ping-pong),
    // Z and the small work buffers on cfg.device.
    explicit Device(const r3d_hip_cfg& cfg) { check(r3d_hip_create(&cfg, &ctx_)); }
    ~Device() { r3d_hip_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;

    // gpu_upload, part 1: global 1-D metrics (lengths npx+2, npy+2, npz+2).
    void set_metrics(const r3d_hip_metrics& m) { check(r3d_hip_set_metrics(ctx_, &m)); }
    // gpu_upload, part 2: one virtual rank's primaries (R3D_RU..R3D_BZ),
    // local arrays with ghosts, x fastest.
    void upload_rank(int rank, const double* const fields[8]) {
        check(r3d_hip_upload_rank(ctx_, rank, fields));
    }
    // gpu_download: fields[R3D_RU..R3D_ZBZ] of one virtual rank (NULL entries
    // are skipped).
    void download_rank(int rank, double* const fields[R3D_NFIELDS]) {
        check(r3d_hip_download_rank(ctx_, rank, fields));
    }
    void set_scalars(const r3d_hip_scalars& s) { check(r3d_hip_set_scalars(ctx_, &s)); }
    r3d_hip_scalars scalars() {
        r3d_hip_scalars s{};
        check(r3d_hip_get_scalars(ctx_, &s));
        return s;
    }
    // step_gpu(...) is step(1); step_gpu(nstages, ...) is step(1, nstages).
    void step(int nsteps = 1, int nstages = 3) { check(r3d_hip_step(ctx_, nsteps, nstages)); }
    // asynchronous checkpoint: begin returns at once; snapshot_rank waits for
    // the copy and returns one rank as download_rank would at begin's point
    void snapshot_begin() { check(r3d_hip_snapshot_begin(ctx_)); }
    r3d_hip_scalars snapshot_rank(int rank, double* const fields[R3D_NFIELDS]) {
        r3d_hip_scalars s{};
        check(r3d_hip_snapshot_rank(ctx_, rank, fields, &s));
        return s;
    }
    void sync() { check(r3d_hip_sync(ctx_)); }
    r3d_hip_ctx* handle() { return ctx_; }

   private:
    r3d_hip_ctx* ctx_ = nullptr;
};

}  // namespace r3d_amd

#endif
