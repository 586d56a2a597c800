// oq_internal.h -- shared host/device definitions of the gfx950 library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/oetqf_amd.h"

namespace oq {

// Thread-local error message behind oq_last_error().
void set_error(const std::string &msg);

// Teardown guard: teardown_sync() waits for the stream (or the device) and
// returns false -- and makes runtime_usable() false for good -- when the
// runtime reports an error, so that destroy functions release nothing
// through a runtime that has been shut down or has faulted.
bool runtime_usable();
bool teardown_sync(hipStream_t s);

#define OQ_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::oq::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return OQ_ERR_HIP;                                                \
        }                                                                     \
    } while (0)

#define OQ_REQUIRE(cond, msg)                \
    do {                                     \
        if (!(cond)) {                       \
            ::oq::set_error(msg);            \
            return OQ_ERR_ARG;               \
        }                                    \
    } while (0)

// Degree trigonometry with exact zeros at multiples of 90 degrees
// (Julia's sind/cosd, used by the reference at mesh.jl:48, GF.jl:79,86).
double sind(double deg);
double cosd(double deg);

// Device copies of the mesh arrays the Green's kernels read.
struct DevFault {
    const double *x, *ax, *y, *z, *axi;
    int64_t nx, nxi;
    double lrept, dep, sd, cd;  // sd, cd: sin/cos of dip
    double s2d, c2d;            // sin/cos of 2*dip (DipSlip traction)
};

struct DevHex8 {
    const double *cx, *cy, *cz, *qx, *qy, *qz, *dx, *dy, *dz;
    int64_t ne;
};

// RAII device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t n);
    int upload(const void *h, size_t n, hipStream_t s = 0);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Green's builders writing element (r, c) of a block to out[r*rs + c*cs]
// (device pointer).  Used by the host-output API (Julia column-major) and
// by oq_rhs_create_from_mesh (dense receiver-major K).
int launch_gf_ff(const DevFault &f, double lambda, double mu, int ftype,
                 int nrept, double *st, hipStream_t s);
int launch_gf_mf(const DevFault &f, const DevHex8 &h, const double *qc,
                 const double *qw, int nq, double lambda, double mu,
                 int ftype, int nrept, double *out, int64_t rs, int64_t cs,
                 hipStream_t s);
int launch_gf_fm(const DevHex8 &h, const DevFault &f, double lambda,
                 double mu, int ftype, double *out, int64_t rs, int64_t cs,
                 hipStream_t s);
int launch_gf_mm(const DevHex8 &h, const double *qc, const double *qw,
                 int nq, double lambda, double mu, double *out, int64_t rs,
                 int64_t cs, hipStream_t s);
// The same pair kernel with receivers from rm and sources from sm: element
// [(k)rm.ne + j, (p)sm.ne + i] at out[row*rs + col*cs].
int launch_gf_mm_pairs(const DevHex8 &rm, const DevHex8 &sm, const double *qc,
                       const double *qw, int nq, double lambda, double mu,
                       double *out, int64_t rs, int64_t cs, hipStream_t s);
// Expands the Toeplitz kernel st (nx,nxi,nxi) into the fault-fault block of
// a dense receiver-major operator (row pitch ld).
int launch_expand_toeplitz(const double *st, int64_t nx, int64_t nxi,
                           double *k, int64_t ld, hipStream_t s);

// Real DFT of length 2nx-1 of the symmetric extension of every (j, l) column
// of the Toeplitz kernel (GF.jl:60-67) and its inverse, on the device:
// st (nx*nxi*nxi) <-> interleaved complex (2*nx*nxi*nxi).
static constexpr int64_t kMaxDftNx = 5120;  // column + twiddles in 160 KB of LDS
int launch_toeplitz_dft(const double *st, int64_t nx, int64_t nxi, double *out, hipStream_t s);
int launch_toeplitz_idft(const double *dft, int64_t nx, int64_t nxi, double *st, hipStream_t s);

// Mesh upload helpers (owning buffers kept by the caller).
struct FaultUpload {
    DevBuf x, ax, y, z, axi;
    DevFault dev{};
    int load(const oq_fault_mesh *m);
};
struct Hex8Upload {
    DevBuf cx, cy, cz, qx, qy, qz, dx, dy, dz;
    DevHex8 dev{};
    int load(const oq_hex8_mesh *m);
};
struct QuadUpload {
    DevBuf c, w;
    int nq = 0;
    int load(const oq_quadrature *q);
};

int check_fault(const oq_fault_mesh *m);
int check_hex8(const oq_hex8_mesh *m);
int check_quad(const oq_quadrature *q);

}  // namespace oq
