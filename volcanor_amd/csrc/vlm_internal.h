// Internal declarations of libvlm (MI355X free-wake hot path).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "vlm_c.h"

namespace vlm {

// Device buffer owned by the library.  reserve() grows and discards the
// contents; growth is geometric (at least doubling) because buffers sized
// by the wake grow every step and hipFree synchronises the device.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t reserve(size_t k)
    {
        if (k <= n && p) return hipSuccess;
        size_t want = k == 0 ? 1 : k;
        if (p && want < 2 * n) want = 2 * n;
        release();
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), want * sizeof(T));
        if (e == hipSuccess) n = want;
        else p = nullptr;
        return e;
    }
};

// Compacted straight-segment list: structure of arrays (stride cap).
//   d64 (9):  ax ay az g | r0x r0y r0z q^2 | n0   g = gam/(4 pi), r0 = b - a,
//                                                 q = rc^2 |r0|^2, n0 = |r0|^2
//   d32 (12): ahx ahy ahz g | alx aly alz q | r0x r0y r0z 4|r0|^2
//             (a = a_hi + a_lo split into two floats)
// count: device int64 with the number of valid segments.
constexpr int SEG64_N = 9;   // doubles per FP64 segment record
constexpr int SEG32_N = 12;  // floats per FP32 segment record

struct SegList {
    DBuf<double> d64;
    DBuf<float> d32;
    DBuf<long long> count;
    int64_t cap = 0;
};

// Staging slots written by the segment builders before compaction.
struct Slot {
    double ax, ay, az, bx, by, bz, gam, rc;
};


// velocity arrays per rotor: the four of include/vlm_c.h, then two more
// histories used only by AM3 and AM4
constexpr int VEL_H2 = 4, VEL_H3 = 5, NVEL = 6;

constexpr int PF_N = 240;  // prescribed far-wake filaments per blade (classdef.f90:225)

// Records of vlm_rotor_set_panels and vlm_rotor_loads (include/vlm_c.h).
constexpr int PANEL_N = 15, SECTION_N = 16, BLADE_N = 9, LOADS_SEC_N = 19, LOADS_ROTOR_N = 9;

struct Rotor {
    vlm_rotor_desc d{};
    int rowNear = 0, rowFar = 0;  // 1-based, as in the reference
    int head = 0;                 // physical slot of logical near-wake row 0
    double shaft[3] = {0, 0, 1}, hub[3] = {0, 0, 0};
    double init_wake_vel = 0.0;   // initWakeVel, applied while iter < init_wake_vel_nt
    int init_wake_vel_nt = 0;
    int nring = 0;                // nb*nc*ns wing rings
    // wing (uploaded by the host after every rigid motion)
    DBuf<double> wc, wrc, wg;     // 12*nring, 4*nring, nring
    // near wake, physical rows
    DBuf<double> node, pnode;     // 3*nb*nN*(ns+1)
    DBuf<double> le, ple;         // 3*nb*(ns+1)
    DBuf<double> gam, rc4, age;   // nb*nN*ns, 4*nb*nN*ns, nb*nN*ns
    // far wake
    DBuf<double> ffc, pffc;       // 6*nb*nF
    DBuf<double> fgam, frc, frc0, fl0, flc, fage;  // nb*nF
    DBuf<double> ftmp;            // shift staging, 11*nb*nF
    // velocities, logical rows: [VLM_VEL, PREV, PRED, STEP] and the older
    // histories of AM3/AM4 (velNwake2, velNwake3; PREV is velNwake1)
    DBuf<double> velN[NVEL], velF[NVEL];
    // prescribed far wake, per wake type (0 current, 1 predicted): PF_N
    // filaments per blade, end points pf[6*(b*PF_N + i) + (0..2 TE, 3..5 LE)],
    // and per blade ps[4*b + (0 gam, 1 core radius, 2 helix pitch, 3 radius)]
    DBuf<double> pf[2], ps[2];
    int presc_nt = 0, presc_gen_nt = 0;  // 0: no prescribed far wake
    double presc_dpsi = 0.0;             // omegaSlow * dt
    int presc_built[2] = {0, 0};
    int iter = 0;                        // step of the last vlm_rotor_begin_step
    // image rotor: index of the rotor it mirrors (-1: a lifting rotor of its
    // own) and the coordinate negated by the mirror (0 x, 1 y, 2 z)
    int image_of = -1, image_axis = 0;
    // right-hand side and loads on the device (vlm_rotor_set_panels,
    // vlm_rotor_rhs, vlm_rotor_loads): panel, section and blade geometry
    // after the last rigid motion (PANEL_N, SECTION_N, BLADE_N doubles
    // each), collocation points SoA, velCP and velCPm (6 per panel), the
    // RHS, the load velocities (SoA), gamPrev (restart state) and the
    // packed load results
    DBuf<double> pan, sec, bla, cps, vcp, rhs, vld, gprev, lout;
    int panels_set = 0;
    size_t nvn() const { return 3u * d.nb * (size_t)d.nNwake * (d.ns + 1); }
    size_t nvf() const { return 3u * d.nb * (size_t)(d.nFwake > 0 ? d.nFwake : 0); }
};

// One profiled Biot-Savart launch: events around the work-queue kernel
// (a, b; null for the small-target kernel, which is counted but not timed)
// and the target count; pairs = targets x the device segment count, which
// is copied to the pinned slot of the same index without a synchronisation.
struct Timer {
    hipEvent_t a, b;
    double pairs;
    int kind = 0;  // 0 work-queue kernel (timed), 1 small-target kernel
};

struct Mlp;

// Learned pass-2 predictor settings, statistics and training buffer.
struct Predictor {
    int mode = 0;          // 0 full Biot-Savart pass 2, 1 learned
    double tol = 1e-4;     // judge tolerance: RMS error / RMS speed on the sample
    int stride = 32;       // judge sample: every stride-th target
    int warmup = 8;        // first AM2 steps always run the full pass
    int train_iters = 4;   // Adam iterations per step on the buffer
    float lr = 1e-3f;
    double lref = 1.0;     // length scale of the neighbour offsets
    int64_t batch = 2048;  // minibatch of each Adam iteration
    int64_t steps = 0, full = 0, learned = 0, fallbacks = 0, judged = 0, train_calls = 0;
    double last_err = 0.0;
    // accumulated over judged steps: squared error of the network, of the
    // network-free substitute, and squared exact speed
    double sum_err2 = 0.0, sum_base2 = 0.0, sum_ref2 = 0.0;
    int64_t buf_cap = 16384, buf_n = 0, buf_head = 0;
    DBuf<float> X, Y, bx, bt, tb;  // tb: minibatch targets in units of the buffer RMS
    DBuf<double> vn, ts, vs, e2, scal;
};

struct FSet {
    DBuf<double> d64;
    DBuf<float> d32;
    DBuf<long long> count;
    int64_t n = 0;
};

struct PSet {
    DBuf<double> p;   // SoA 3*m
    DBuf<double> v;   // SoA 3*m
    int64_t m = 0;
};

}  // namespace vlm

struct vlm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = VLM_FP64;
    int bug_compat = 1;
    std::string err;
    std::vector<vlm::Rotor *> rotors;
    std::vector<vlm::FSet *> fsets;
    std::vector<vlm::PSet *> psets;
    // scratch
    vlm::SegList seg;
    vlm::DBuf<vlm::Slot> slots;
    vlm::DBuf<int> flags, idx;
    vlm::DBuf<unsigned char> scan_tmp;
    size_t scan_tmp_bytes = 0;
    vlm::DBuf<double> tgt, out, part;
    vlm::DBuf<int> counter;  // run counts per workgroup of the Biot-Savart launches
    vlm::DBuf<int4> skruns;  // run lists (tile, first batch, batch count) per workgroup
    int occ[4] = {0, 0, 0, 0};  // resident workgroups per kernel variant
    vlm::DBuf<double> hostio;  // device staging for host point lists
    // pinned host arena for host<->device copies: truly asynchronous copies,
    // reset whenever the stream has been synchronised
    char *pin = nullptr;
    size_t pin_cap = 0, pin_off = 0;
    // profiling
    int profile = 0;
    std::vector<vlm::Timer> timers;
    std::vector<vlm::Timer> lu_timers;  // LU factorisations (pairs = flops)
    double prof_ms = 0.0, prof_pairs = 0.0;
    long long *prof_nseg = nullptr;  // pinned, one slot per timed launch
    size_t prof_cap = 0;
    int64_t prof_launches = 0;
    // dense solve state per rotor (AIC, LU, pivots)
    std::vector<double *> aic, lu;
    std::vector<int *> piv;
    std::vector<int> aic_n;
    vlm::DBuf<double> solve_tmp;
    vlm::DBuf<int> lu_work;   // row-move lists of the blocked LU
    vlm::Mlp *mlp = nullptr;
    vlm::Predictor pred;
    hipEvent_t judge_ev = nullptr;  // learned pass: the judge sums reached judge_pin
    double *judge_pin = nullptr;    // pinned, 4 doubles
    // deferred downloads: device -> pinned arena copies already enqueued,
    // copied to the caller's buffers at vlm_wait or at any stream sync
    struct Deferred {
        void *dst;
        const char *src;
        size_t bytes;
    };
    std::vector<Deferred> deferred;
    hipEvent_t defer_ev = nullptr;  // recorded after the last deferred copy
};

namespace vlm {

int fail(vlm_ctx *ctx, int code, const std::string &msg);
int hip_fail(vlm_ctx *ctx, hipError_t e, const char *what);

#define VLM_HIP(call)                                          \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return vlm::hip_fail(ctx, e_, #call); \
    } while (0)

// Every entry point that touches the GPU makes the context's device current.
#define VLM_DEVICE(ctx) VLM_HIP(hipSetDevice((ctx)->device))

#define VLM_TRY(call)                 \
    do {                              \
        int rc_ = (call);             \
        if (rc_ != VLM_OK) return rc_; \
    } while (0)

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kInv4Pi = 0.25 / kPi;
constexpr double kEps = 2.220446049250313080847263336181640625e-16;  // epsilon(1._dp)

// The FP64 and FP32 records of one segment at position k of a compacted
// list of stride cap (layouts above).
__device__ __forceinline__ void write_segment_record(const Slot &s, int64_t k, double *d64,
                                                     float *d32, int64_t cap)
{
    const double g = s.gam * kInv4Pi;
    const double r0x = s.bx - s.ax, r0y = s.by - s.ay, r0z = s.bz - s.az;
    const double n0 = r0x * r0x + r0y * r0y + r0z * r0z;
    const double q = s.rc * s.rc * n0;
    d64[k] = s.ax;
    d64[cap + k] = s.ay;
    d64[2 * cap + k] = s.az;
    d64[3 * cap + k] = g;
    d64[4 * cap + k] = r0x;
    d64[5 * cap + k] = r0y;
    d64[6 * cap + k] = r0z;
    d64[7 * cap + k] = q * q;
    d64[8 * cap + k] = n0;
    const float hx = (float)s.ax, hy = (float)s.ay, hz = (float)s.az;
    d32[k] = hx;
    d32[cap + k] = hy;
    d32[2 * cap + k] = hz;
    d32[3 * cap + k] = (float)g;
    d32[4 * cap + k] = (float)(s.ax - (double)hx);
    d32[5 * cap + k] = (float)(s.ay - (double)hy);
    d32[6 * cap + k] = (float)(s.az - (double)hz);
    d32[7 * cap + k] = (float)q;
    d32[8 * cap + k] = (float)r0x;
    d32[9 * cap + k] = (float)r0y;
    d32[10 * cap + k] = (float)r0z;
    d32[11 * cap + k] = (float)(4.0 * n0);
}

// Biot-Savart evaluation of a compacted segment list at n targets stored
// as SoA (x at t[0..], y at t[tstride..], z at t[2*tstride..]).  Result
// goes to out (SoA, same stride); accumulate adds into it.  nseg_max is an
// upper bound of the device count used to size the launch.
// A fixed workspace for bs_eval instead of the context's growing buffers
// (no allocation): room for part_cap doubles of chunk partial sums (at
// least 3 m, the chunk count is capped by it) and one int counter.  Such
// launches are not profiled.
struct BsWork {
    double *part;
    size_t part_cap;
    int *counter;
};
int bs_eval(vlm_ctx *ctx, const double *seg64, const float *seg32,
            int64_t segstride, const long long *nseg_dev, int64_t nseg_max,
            const double *t, int64_t m, int64_t tstride, double *out,
            int accumulate, const BsWork *ws = nullptr);

// Compact staged slots (flags != 0) into ctx->seg.  nslots on host.
int compact_slots(vlm_ctx *ctx, int64_t nslots);
int compact_into(vlm_ctx *ctx, const Slot *slots, const int *flags,
                 int64_t nslots, DBuf<double> &d64, DBuf<float> &d32,
                 DBuf<long long> &count, int64_t cap);

// Source-set builders write slots starting at slot offset off; each
// returns the number of slots it used through *used.
int stage_wing(vlm_ctx *ctx, Rotor &r, int set, int64_t off, int64_t *used);
int stage_wake(vlm_ctx *ctx, Rotor &r, int pred, int64_t off, int64_t *used);
int64_t wing_slots(const Rotor &r, int set);
int64_t wake_slots(const Rotor &r);

int ensure_slots(vlm_ctx *ctx, int64_t n);

// Rotor ir of the context (nullptr when out of range).
Rotor *rotor_of(vlm_ctx *ctx, int ir);
// The ordered sum of source-set terms (vlm_induced_velocity_terms) at m
// device targets, SoA with stride m, into the device SoA buffer out.
int induced_terms_soa(vlm_ctx *ctx, int nterms, const int *ir, const int *set,
                      const int *flags, int64_t m, const double *tgt, double *out);
void mlp_release(vlm_ctx *ctx);

// Host<->device copies through the pinned arena (ctx->pin): h2d copies the
// host bytes into the arena and enqueues an asynchronous copy, so the
// caller's buffer is free on return without a stream synchronisation;
// sync_stream waits for the stream and recycles the arena; d2h enqueues the
// copy, synchronises and copies out of the arena.
int pin_take(vlm_ctx *ctx, size_t bytes, char **out);
int h2d(vlm_ctx *ctx, void *dst, const void *src, size_t bytes);
int d2h(vlm_ctx *ctx, void *dst, const void *src, size_t bytes);
int sync_stream(vlm_ctx *ctx);
// Deferred download: the copy is enqueued now, dst is written at
// wait_deferred or at the next sync_stream (whichever comes first).
int d2h_deferred(vlm_ctx *ctx, void *dst, const void *src, size_t bytes);
int wait_deferred(vlm_ctx *ctx);
int mlp_forward_dev(vlm_ctx *ctx, int64_t m, const float *X, float *Y);
// Training step on a device batch; loss_host downloads the loss (and
// synchronises), keep_loss only sums it on the device for mlp_read_loss.
int mlp_train_dev(vlm_ctx *ctx, int64_t m, const float *X, const float *T, int adam, float lr,
                  float *loss_host, int keep_loss = 0);
int mlp_read_loss(vlm_ctx *ctx, float *loss);

}  // namespace vlm
