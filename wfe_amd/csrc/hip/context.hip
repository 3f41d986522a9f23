// Integrator context for gfx950: device buffers, the RK3/acoustic schedule
// and the C ABI of include/wfe_hip.h (dynamics part).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "wfe_hip.h"
#include "hip/acoustic_kernels.hpp"
#include "hip/kessler_kernels.hpp"
#include "hip/slow_kernels.hpp"
#include "hip/adjoint_kernels.hpp"
#include "hip/output_kernels.hpp"
#include "hip/surface_kernels.hpp"

namespace wfe {

thread_local std::string g_last_error;

void set_last_error(const std::string& s) { g_last_error = s; }

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

namespace {

#define WFE_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::wfe::HipError(std::string("HIP error ") + hipGetErrorName(e_) + " at " __FILE__ \
                     ":" + std::to_string(__LINE__) + ": " + hipGetErrorString(e_));      \
  } while (0)

// WFE_SYNC=1 (the reference's launch-check switch, src/core/cuda_check.hpp:
// 20-32): every launch check also waits for the device, so an asynchronous
// fault is reported at the kernel that caused it; step graphs are then off.
// Launches recorded into a graph are not waited for (capture_depth > 0).
bool sync_launches() {
  static const bool on = [] {
    const char* e = std::getenv("WFE_SYNC");
    return e && e[0] == '1';
  }();
  return on;
}
thread_local int capture_depth = 0;

struct CaptureScope {
  CaptureScope() { ++capture_depth; }
  ~CaptureScope() { --capture_depth; }
};

inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && sync_launches() && capture_depth == 0) e = hipDeviceSynchronize();
  if (e != hipSuccess)
    throw HipError(std::string("kernel error in ") + what + ": " + hipGetErrorString(e));
}

inline unsigned cdiv(int a, int b) { return (unsigned)((a + b - 1) / b); }

// Owns every device allocation of a context (zero-initialised).
class DevArena {
 public:
  ~DevArena() {
    for (void* p : ptrs_) (void)hipFree(p);
  }
  template <class E>
  E* alloc(size_t n) {
    void* p = nullptr;
    WFE_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(E)));
    ptrs_.push_back(p);
    WFE_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(E)));
    bytes_ += n * sizeof(E);
    return static_cast<E*>(p);
  }
  size_t bytes() const { return bytes_; }

 private:
  std::vector<void*> ptrs_;
  size_t bytes_ = 0;
};

template <class T>
struct StateSet {
  T *u = nullptr, *v = nullptr, *w = nullptr, *thp = nullptr;
  T *qv = nullptr, *qc = nullptr, *qr = nullptr;
  double* pip = nullptr;
};

// Optional per-launch timing with HIP events recorded on the context's
// stream around each kernel (wfe_kernel_timing). Event pairs are collected
// lazily; totals are per kernel name.
class KernelTimer {
 public:
  ~KernelTimer() {
    for (auto& r : pending_) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }
  bool on = false;
  void begin(const char* name, hipStream_t s) {
    if (!on) return;
    cur_ = Rec{name, get(), get()};
    WFE_HIP(hipEventRecord(cur_.a, s));
  }
  void end(hipStream_t s) {
    if (!on) return;
    WFE_HIP(hipEventRecord(cur_.b, s));
    pending_.push_back(cur_);
    if (pending_.size() > 4096) collect();
  }
  void collect() {
    for (auto& r : pending_) {
      WFE_HIP(hipEventSynchronize(r.b));
      float ms = 0;
      WFE_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      auto& a = acc_[r.name];
      a.first += 1;
      a.second += ms;
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    pending_.clear();
  }
  std::string report() {
    collect();
    std::string out;
    char line[256];
    for (auto& kv : acc_) {
      std::snprintf(line, sizeof(line), "%s %ld %.6f\n", kv.first.c_str(), kv.second.first,
                    kv.second.second);
      out += line;
    }
    return out;
  }
  void reset() {
    collect();
    acc_.clear();
  }

 private:
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  hipEvent_t get() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    WFE_HIP(hipEventCreate(&e));
    return e;
  }
  Rec cur_{};
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> pool_;
  std::map<std::string, std::pair<long, double>> acc_;
};

// Launch through the per-kernel timer; when a solo measurement of `name`
// is armed (kernel_bench), the launch is repeated back to back between two
// events instead.
#define TIMED(kname_, ...)                                      \
  do {                                                          \
    if (solo_.name && std::strcmp(solo_.name, kname_) == 0) {   \
      solo_.run(stream_, [&] { __VA_ARGS__; });                 \
    } else {                                                    \
      timer_.begin(kname_, stream_);                            \
      __VA_ARGS__;                                              \
      timer_.end(stream_);                                      \
    }                                                           \
  } while (0)

// Back-to-back timing of one kernel with its arguments of the moment.
struct SoloBench {
  const char* name = nullptr;
  int reps = 0;
  double avg_us = -1;
  template <class F>
  void run(hipStream_t s, F&& launch) {
    hipEvent_t e0, e1;
    WFE_HIP(hipEventCreate(&e0));
    WFE_HIP(hipEventCreate(&e1));
    for (int r = 0; r < 3; ++r) launch();
    WFE_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) launch();
    WFE_HIP(hipEventRecord(e1, s));
    WFE_HIP(hipEventSynchronize(e1));
    float ms = 0;
    WFE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    avg_us = 1e3 * ms / reps;
  }
};

}  // namespace

// Type-erased interface used by the C ABI.
struct CtxBase {
  virtual ~CtxBase() = default;
  // One large step; t is the time at its start (the surface physics uses
  // the hour of day at t + dt).
  virtual void step(double dt, double t) = 0;
  virtual void set_physics(const wfe_sfc& sfc) = 0;
  virtual void upload(int field, const double* h, size_t n) = 0;
  virtual void download(int field, double* h, size_t n) = 0;
  virtual float absmax(int field) = 0;
  // Interior diagnostics record (dev::kDg* slots) of the time-n state.
  virtual void diagnose(double* rec) = 0;
  // Interior of a state field (or the rain) packed as float32 in output
  // order and copied asynchronously into dst; output_mark returns a ticket
  // for the copies enqueued so far, output_wait blocks until they are done.
  virtual void pack_async(int field, float* dst, size_t n) = 0;
  virtual long long output_mark() = 0;
  virtual void output_wait(long long ticket) = 0;
  virtual void sync() = 0;
  virtual void set_bdy_weights(const double* w, size_t n) = 0;
  virtual void set_bdy_target(int slot, int var, const double* h, size_t n) = 0;
  virtual void swap_bdy_targets() = 0;
  virtual void set_bdy_tfrac(double tf) = 0;
  virtual void apply_bcs() = 0;
  virtual void compute_tendencies(double dt) = 0;
  virtual void acoustic_substeps(double dtau, int nsub) = 0;
  virtual void kessler(double dt) = 0;
  virtual size_t device_bytes() const = 0;
  virtual size_t host_n3() const = 0;  // reference-layout element counts
  virtual size_t host_n2() const = 0;
  virtual void kernel_timing(int enable) = 0;
  virtual std::string kernel_report() = 0;
  virtual double kernel_bench(const std::string& name, int reps) = 0;
  // Ensemble support: the context's stream, whether its first (tuning) step
  // has run, the parity of its state buffers, the generation of everything
  // baked into its kernel arguments, and the step's launches enqueued on
  // its stream without changing the parity (for graph capture).
  virtual hipStream_t stream() const = 0;
  virtual bool ready_for_capture() const = 0;
  virtual int parity() const = 0;
  virtual void flip_parity() = 0;
  virtual unsigned generation() const = 0;
  virtual void enqueue_step(double dt) = 0;
  virtual bool surface_physics() const = 0;
  // The surface physics' hour of day for a step from t to t + dt, written
  // on the context's stream (a device scalar, so step graphs follow it).
  virtual void set_clock(double dt, double t) = 0;
};

template <class T>
class Adjoint;

template <class T>
class Ctx final : public CtxBase {
  friend class Adjoint<T>;
 public:
  Ctx(const wfe_grid& gr, const wfe_params& hp, const wfe_base& base, const wfe_metric& met);
  ~Ctx() override {
    for (hipGraphExec_t& e : gexec_)
      if (e) (void)hipGraphExecDestroy(e);
    if (ostream_) {
      (void)hipStreamSynchronize(ostream_);
      (void)hipStreamDestroy(ostream_);
    }
    for (hipEvent_t& e : out_packed_)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t& e : out_copied_)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t& e : out_ring_)
      if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  void step(double dt, double t) override;
  void set_clock(double dt, double t) override;
  void set_physics(const wfe_sfc& sfc) override;
  void upload(int field, const double* h, size_t n) override;
  void download(int field, double* h, size_t n) override;
  float absmax(int field) override;
  void diagnose(double* rec) override;
  void pack_async(int field, float* dst, size_t n) override;
  long long output_mark() override;
  void output_wait(long long ticket) override;
  void sync() override { WFE_HIP(hipStreamSynchronize(stream_)); }
  void set_bdy_weights(const double* w, size_t n) override;
  void set_bdy_target(int slot, int var, const double* h, size_t n) override;
  void swap_bdy_targets() override {
    std::swap(bdy_lo_, bdy_hi_);
    ++gen_;
  }
  void set_bdy_tfrac(double tf) override {
    tfrac_ = tf;
    dev::k_set_scalar<T><<<1, 64, 0, stream_>>>(tf_dev_, (T)tf);
    check_launch("set_bdy_tfrac");
  }
  void apply_bcs() override { fill_ghosts(S_[cur_]); }
  void compute_tendencies(double dt) override { slow_tendencies(S_[cur_], dt); }
  void acoustic_substeps(double dtau, int nsub) override;
  void kessler(double dt) override { kessler_step(S_[cur_], dt); }
  size_t device_bytes() const override { return arena_.bytes(); }
  size_t host_n3() const override { return hn3_; }
  size_t host_n2() const override { return hn2_; }
  void kernel_timing(int enable) override {
    WFE_HIP(hipStreamSynchronize(stream_));
    if (enable == 2) timer_.reset();
    timer_.on = enable != 0;
  }
  double kernel_bench(const std::string& name, int reps) override;
  hipStream_t stream() const override { return stream_; }
  bool ready_for_capture() const override {
    const bool kessler = p_.moisture && p_.microphysics && g_.nz <= 48;
    return tuned_ && tend_tuned_ && (kes_tuned_ || !kessler) && !timer_.on;
  }
  int parity() const override { return cur_; }
  void flip_parity() override { cur_ ^= 1; }
  unsigned generation() const override { return gen_; }
  void enqueue_step(double dt) override {
    const int par = cur_;
    step_launches(dt);
    cur_ = par;
  }

  // The pieces of a step (step_launches), for the member-batched ensemble
  // step. Stage m reads stage_in(m) and writes stage_out(m); the time-n
  // state is S_[cur_].
  int stage_nsub(int m) const {
    const int ns = std::max(1, hp_.acoustic_ns);
    return m == 0 ? 1 : m == 1 ? std::max(1, ns / 2) : ns;
  }
  static double stage_dt(double dt, int m) { return m == 0 ? dt / 3 : m == 1 ? dt / 2 : dt; }
  const StateSet<T>& stage_in(int m) const { return m == 0 ? S_[cur_] : SE_[m - 1]; }
  const StateSet<T>& stage_out(int m) const { return m == 2 ? S_[cur_ ^ 1] : SE_[m]; }
  // Without boundary relaxation (which adds to T_q after the tendencies)
  // the moisture stage update is applied where T_q is formed.
  bool fuse_q() const { return p_.moisture && !bdy_active_; }
  // Ghost fill, slow tendencies and relaxation of stage m.
  void stage_pre(double dt, int m) {
    const StateSet<T>& est = stage_in(m);
    fill_ghosts(est);
    slow_tendencies(est, dt, fuse_q() ? &stage_out(m) : nullptr, stage_dt(dt, m), true);
    if (bdy_active_) bdy_relax(est);
  }
  // The moisture stage update of stage m where it is not fused.
  void stage_post(double dt, int m) {
    if (p_.moisture && !fuse_q()) stage_moisture(stage_out(m), stage_dt(dt, m));
  }
  // Kessler and the surface physics on the new state S_[cur_].
  void step_post(double dt) {
    if (p_.moisture && p_.microphysics) kessler_step(S_[cur_], dt);
    if (phys_on_) surface_step(S_[cur_], dt);  // integrator.cpp:109
  }
  // This member's entry for substep s of stage m in a batched launch.
  void batch_member(dev::BatchMember<T>& M, double dt, int m, int s) const;
  int uv_tile_width() const { return lv_bx_; }
  const dev::Grid& grid() const { return g_; }
  int levels() const { return g_.nz; }
  bool relaxing() const { return bdy_active_; }
  bool moist() const { return p_.moisture != 0; }
  bool surface_physics() const override { return phys_on_; }
  std::string kernel_report() override {
    WFE_HIP(hipStreamSynchronize(stream_));
    std::string r = timer_.report();
    if (tuned_) r += "tuned:" + methods_[method_].name + " 0 0\n";
    r += std::string("uniform_columns:") + (p_.hom ? "1" : "0") + " 0 0\n";
    r += tune_report_;
    r += tend_report_;
    r += kes_report_;
    return r;
  }

 private:
  // host helpers
  enum class Shape { k3D, k2D, k1D };  // padded 3-D, padded plane, NZ column
  template <class E>
  E* upload_new(const double* h, Shape shape);
  void fill_ghosts(const StateSet<T>& s);
  void slow_tendencies(const StateSet<T>& est, double dt, const StateSet<T>* qout = nullptr,
                       double dt_stage = 0, bool fold_buoyancy = false);
  void bdy_relax(const StateSet<T>& est);
  void acoustic_stage(const StateSet<T>& out, double dtau, int nsub);
  dev::FusedArgs<T> substep_args(const StateSet<T>& out, double dtau, int s, int nsub) const;
  void stage_moisture(const StateSet<T>& out, double dt_stage);
  void kessler_step(const StateSet<T>& s, double dt);
  T* field_ptr(int field);
  dim3 blk3() const { return dim3(64, 4, 1); }
  // Average time [ms] of `f` (a launch sequence on the stream) over 5 runs
  // after 2 warm-up runs, with the per-kernel timer paused.
  template <class F>
  float time_launches(F&& f) {
    hipEvent_t e0, e1;
    WFE_HIP(hipEventCreate(&e0));
    WFE_HIP(hipEventCreate(&e1));
    const bool was_on = timer_.on;
    timer_.on = false;
    for (int r = 0; r < 2; ++r) f();
    WFE_HIP(hipEventRecord(e0, stream_));
    for (int r = 0; r < 5; ++r) f();
    WFE_HIP(hipEventRecord(e1, stream_));
    WFE_HIP(hipEventSynchronize(e1));
    timer_.on = was_on;
    float ms = 0;
    WFE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_launch("tuning");
    return ms / 5;
  }
  void launch_uv_tile(const dev::UVArgs<T>& uv) {
    const dim3 guv = lv_grid(g_.nx + 1, g_.ny + 1, g_.nz);
    if (lv_bx_ == 32)
      TIMED("acou_uv", (dev::k_acou_uv_tile<T, 32><<<guv, lv_block(), 0, stream_>>>(
                           g_, p_, base_, met_, uv)));
    else
      TIMED("acou_uv", (dev::k_acou_uv_tile<T, 64><<<guv, lv_block(), 0, stream_>>>(
                           g_, p_, base_, met_, uv)));
  }
  // Workgroups of the step's per-level kernels: 256 threads, lv_bx_ wide
  // (64, or 32 where 32-wide tiles leave fewer idle lanes along x).
  int lv_bx_ = 64;
  dim3 lv_block() const { return dim3(lv_bx_, 256 / lv_bx_, 1); }
  dim3 lv_grid(int ni, int nj, int nk) const {
    return dim3(cdiv(ni, lv_bx_), cdiv(nj, 256 / lv_bx_), (unsigned)nk);
  }

  dev::Grid g_{};
  dev::Params p_{};
  wfe_params hp_{};
  size_t npts_ = 0, n2_ = 0;     // device-layout element counts (3-D, 2-D)
  int hNX_ = 0;                   // reference-layout row length nx + 2 ng
  size_t hn3_ = 0, hn2_ = 0;      // reference-layout element counts
  template <class E>
  std::vector<E> to_dev(const double* h, bool three_d) const;
  template <class E>
  void to_host(const std::vector<E>& d, double* h, bool three_d) const;
  hipStream_t stream_ = nullptr;
  DevArena arena_;
  dev::Base<T> base_{};
  dev::Metric<T> met_{};
  dev::Coef<T> coef_{};
  StateSet<T> S_[2]{};
  StateSet<T> SE_[2]{};  // stage estimates (ping-pong: a stage never writes what it reads)
  int cur_ = 0;
  T* wu_[2]{};
  T* wv_[2]{};
  T* ww_ = nullptr;
  T *tu_ = nullptr, *tv_ = nullptr, *tw_ = nullptr, *tth_ = nullptr, *tpi_ = nullptr;
  // T_w with the moist buoyancy for the phase-level substeps (moist cases);
  // tw_sub_ points at it while they run, else the substeps read tw_.
  T* twf_ = nullptr;
  const T* tw_sub_ = nullptr;
  // Levels per thread of the tendency kernel (WFE_TENDENCY_LEVELS: 1, 2, 4;
  // 2 measured fastest on the supercell).
  int tend_levels_ = 2;
  // Columns per workgroup row of the tendency kernel (WFE_TENDENCY_TILE:
  // 64 -> 64x4 threads, 32 -> 32x8; the two measured alike).
  int tend_tx_ = 64;
  // Kessler workgroup shape for columns of at most 48 levels
  // (WFE_KESSLER_TILE: 64x8, 32x8, 32x16, 16x16, 64x16, 32x24, 16x48 columns
  // x thread rows). 32x16 (3 levels per thread, 480 workgroups on the
  // 120x120x40 supercell) measured 18.6 us against 21.8 us for 64x8.
  std::string kes_tile_ = "32x16";
  // Tendency (levels per thread, tile width) and Kessler tile shapes are
  // timed on the first step unless pinned through the environment.
  bool tend_tuned_ = false, kes_tuned_ = false;
  std::string tend_report_, kes_report_;
  T* kes_rain_scratch_ = nullptr;
  T* tq_[3]{};
  T *mfx_ = nullptr, *mfy_ = nullptr, *mfz_ = nullptr, *div_ = nullptr;
  T* pratio_[3]{};
  T* rain_ = nullptr;
  bool bdy_active_ = false;
  T *bdy_lo_[5]{}, *bdy_hi_[5]{};
  T* bdy_wgt_ = nullptr;
  int bdy_rim_ = 0;  // relaxation launched over this rim only (0: whole grid)
  double tfrac_ = 0;
  unsigned* d_absmax_ = nullptr;
  // Surface/PBL column physics after every step (set_physics): FP64 fields
  // and the hour of day at the end of the step as a device scalar, so a
  // captured step graph stays valid.
  bool phys_on_ = false;
  double phys_start_hour_ = 0;
  dev::SfcArgs sfc_{};
  double* utc_dev_ = nullptr;
  void surface_step(const StateSet<T>& s, double dt);
  double* phys_field(int f) const;
  // Diagnostics partial records (one per workgroup of k_diag_partial) and
  // the final record after them.
  double* diag_part_ = nullptr;
  int diag_nb_ = 0;
  // Asynchronous output: a float32 staging array per packed field, filled
  // on stream_ and copied to the host on ostream_ (default priority), with
  // per-field events so a field is not repacked before its last copy ended;
  // out_ring_ holds the tickets of output_mark.
  hipStream_t ostream_ = nullptr;
  float* out_stage_[WFE_F_COUNT]{};
  hipEvent_t out_packed_[WFE_F_COUNT]{};
  hipEvent_t out_copied_[WFE_F_COUNT]{};
  bool out_used_[WFE_F_COUNT]{};
  static constexpr int kOutRing = 64;
  hipEvent_t out_ring_[kOutRing]{};
  long long out_seq_ = 0;
  std::mutex out_mu_;
  // Acoustic substep methods (split u/v + column-solve pairs and fused
  // shapes); all run under the same buffer rotation, the fastest is chosen
  // by a timing pass on the first stage.
  struct SubstepMethod {
    std::string name;
    std::function<void(const dev::FusedArgs<T>&)> run;
  };
  std::vector<SubstepMethod> methods_;
  int method_ = -1;
  bool tuned_ = false;
  // Step graphs, one per parity of cur_ (the two state sets swap roles),
  // re-captured when dt or anything baked into kernel arguments changes
  // (generation gen_: boundary buffers, substep method).
  void step_launches(double dt);
  bool use_graph_ = true;
  hipGraphExec_t gexec_[2] = {nullptr, nullptr};
  double gdt_[2] = {0, 0};
  unsigned ggen_[2] = {0, 0};
  unsigned gen_ = 1;
  T* tf_dev_ = nullptr;
  std::string tune_report_;
  double* xpip_[3]{};  // rotating pi' work buffers
  T* xthp_[2]{};       // ping-pong theta' work buffers
  void build_methods(const hipDeviceProp_t& prop);
  void tune_acoustic(double dtau);
  template <int TX, int NW, int KPT, int WPE = 1>
  void add_split_tile(const char* name, size_t cap);
  template <int TX, int NW, int KPT, int WPE = 1>
  void add_row(const char* name, size_t cap);
  template <int CX, int CY, int NW, int KPT>
  void add_tile2d(const char* name, size_t cap);
  KernelTimer timer_;
  SoloBench solo_;
};

// Reference layout (row length nx + 2ng) <-> device layout (pitch NX, i = 0
// at column xo). Row padding stays zero on the device.
template <class T>
template <class E>
std::vector<E> Ctx<T>::to_dev(const double* h, bool three_d) const {
  const int nk = three_d ? g_.NZ : 1;
  std::vector<E> d(three_d ? npts_ : n2_, (E)0);
  for (int kr = 0; kr < nk; ++kr)
    for (int jr = 0; jr < g_.NY; ++jr) {
      const double* src = h + ((size_t)kr * g_.NY + jr) * hNX_;
      E* dst = d.data() + ((size_t)kr * g_.NY + jr) * g_.NX + (g_.xo - g_.ng);
      for (int ir = 0; ir < hNX_; ++ir) dst[ir] = (E)src[ir];
    }
  return d;
}

template <class T>
template <class E>
void Ctx<T>::to_host(const std::vector<E>& d, double* h, bool three_d) const {
  const int nk = three_d ? g_.NZ : 1;
  for (int kr = 0; kr < nk; ++kr)
    for (int jr = 0; jr < g_.NY; ++jr) {
      const E* src = d.data() + ((size_t)kr * g_.NY + jr) * g_.NX + (g_.xo - g_.ng);
      double* dst = h + ((size_t)kr * g_.NY + jr) * hNX_;
      for (int ir = 0; ir < hNX_; ++ir) dst[ir] = (double)src[ir];
    }
}

// Allocates a device array of the given shape and, when h is non-null,
// uploads h (reference layout, FP64) converted to E; otherwise zeros.
template <class T>
template <class E>
E* Ctx<T>::upload_new(const double* h, Shape shape) {
  const size_t nd = shape == Shape::k3D ? npts_ : (shape == Shape::k2D ? n2_ : (size_t)g_.NZ);
  E* d = arena_.alloc<E>(nd);
  if (h) {
    const std::vector<E> tmp = shape == Shape::k1D ? std::vector<E>(h, h + nd)
                                                   : to_dev<E>(h, shape == Shape::k3D);
    WFE_HIP(hipMemcpy(d, tmp.data(), nd * sizeof(E), hipMemcpyHostToDevice));
  }
  return d;
}

template <class T>
Ctx<T>::Ctx(const wfe_grid& gr, const wfe_params& hp, const wfe_base& b, const wfe_metric& m) {
  hp_ = hp;
  g_.nx = gr.nx;
  g_.ny = gr.ny;
  g_.nz = gr.nz;
  g_.ng = gr.ng;
  g_.xo = dev::kRowAlign;
  g_.NX = (g_.xo + gr.nx + gr.ng + dev::kRowAlign - 1) / dev::kRowAlign * dev::kRowAlign;
  g_.NY = gr.ny + 2 * gr.ng;
  g_.NZ = gr.nz + 1 + 2 * gr.ng;
  g_.plane = g_.NX * g_.NY;
  hNX_ = gr.nx + 2 * gr.ng;
  hn2_ = (size_t)hNX_ * g_.NY;
  hn3_ = hn2_ * g_.NZ;
  g_.dx = gr.dx;
  g_.dy = gr.dy;
  g_.rdx = 1.0 / gr.dx;
  g_.rdy = 1.0 / gr.dy;
  npts_ = (size_t)g_.plane * g_.NZ;
  n2_ = (size_t)g_.plane;
  p_.diff_K = hp.diff_K;
  p_.coriolis_use_ub = hp.coriolis_use_ub;
  p_.rayleigh_zd = hp.rayleigh_zd;
  p_.rayleigh_alpha = hp.rayleigh_alpha;
  p_.beta = hp.acoustic_beta;
  p_.smdiv = hp.acoustic_smdiv;
  p_.bcx = hp.bc_x_open;
  p_.bcy = hp.bc_y_open;
  p_.cstar = hp.cstar;
  p_.moisture = hp.moisture;
  p_.pd_moist = hp.moisture && hp.pd_moist;
  p_.w_damping = hp.w_damping;
  p_.microphysics = hp.microphysics;

  if (const char* e = std::getenv("WFE_GRAPH")) use_graph_ = std::string(e) != "0";
  if (sync_launches()) use_graph_ = false;
  if (const char* e = std::getenv("WFE_TENDENCY_LEVELS")) {
    const int v = std::atoi(e);
    if (v != 1 && v != 2 && v != 4) throw ArgError("WFE_TENDENCY_LEVELS must be 1, 2 or 4");
    tend_levels_ = v;
    tend_tuned_ = true;
  }
  lv_bx_ = cdiv(gr.nx, 32) * 32 < cdiv(gr.nx, 64) * 64 ? 32 : 64;
  tend_tx_ = lv_bx_;
  if (const char* e = std::getenv("WFE_TENDENCY_TILE")) {
    const int v = std::atoi(e);
    if (v != 64 && v != 32) throw ArgError("WFE_TENDENCY_TILE must be 64 or 32");
    tend_tx_ = v;
    tend_tuned_ = true;
  }
  if (const char* e = std::getenv("WFE_KESSLER_TILE")) {
    const std::string v = e;
    if (v != "64x8" && v != "32x8" && v != "32x16" && v != "16x16" && v != "64x16" &&
        v != "32x24" && v != "16x48")
      throw ArgError("WFE_KESSLER_TILE must be 64x8, 32x8, 32x16, 16x16, 64x16, 32x24 or 16x48");
    kes_tile_ = v;
    kes_tuned_ = true;
  }
  WFE_HIP(hipStreamCreate(&stream_));
  const size_t n = npts_;
  const size_t hn = hn3_;
  const Shape s3 = Shape::k3D, s2 = Shape::k2D, s1 = Shape::k1D;

  // Base state: FP64 host arrays -> T (pibar stays FP64).
  base_.thb = upload_new<T>(b.thb, s3);
  base_.rhob = upload_new<T>(b.rhob, s3);
  base_.dthbdz = upload_new<T>(b.dthbdz, s3);
  base_.thbw = upload_new<T>(b.thbw, s3);
  base_.rhobw = upload_new<T>(b.rhobw, s3);
  base_.qvb = upload_new<T>(b.qvb, s3);
  base_.ub = upload_new<T>(b.ub, s1);
  base_.vb = upload_new<T>(b.vb, s1);  // NULL -> zeros
  base_.pib = upload_new<double>(b.pib, s3);

  met_.zeta_w = upload_new<T>(m.zeta_w, s1);
  met_.zeta_c = upload_new<T>(m.zeta_c, s1);
  met_.dzeta_c = upload_new<T>(m.dzeta_c, s1);
  met_.dzeta_w = upload_new<T>(m.dzeta_w, s1);
  met_.hx = upload_new<T>(m.hx, s2);
  met_.hy = upload_new<T>(m.hy, s2);
  met_.hxu = upload_new<T>(m.hx_u, s2);
  met_.hyv = upload_new<T>(m.hy_v, s2);
  met_.jac = upload_new<T>(m.jac, s2);
  met_.fcor = upload_new<T>(m.fcor, s2);
  met_.mapf = upload_new<T>(m.mapf, s2);
  met_.zt = (T)m.zt;
  p_.terrain = 0;
  for (size_t q = 0; q < hn2_; ++q)
    if (m.hx[q] != 0 || m.hy[q] != 0 || m.hx_u[q] != 0 || m.hy_v[q] != 0) p_.terrain = 1;
  met_.rzt = 1.0 / m.zt;
  {
    std::vector<double> rc(g_.NZ), rw(g_.NZ);
    for (int k = 0; k < g_.NZ; ++k) {
      rc[k] = m.dzeta_c[k] != 0 ? 1.0 / m.dzeta_c[k] : 0.0;
      rw[k] = m.dzeta_w[k] != 0 ? 1.0 / m.dzeta_w[k] : 0.0;
    }
    met_.rdzc = upload_new<double>(rc.data(), s1);
    met_.rdzw = upload_new<double>(rw.data(), s1);
    // sin^2 Rayleigh layer above rayleigh_zd (ray_alpha, kernels.cu:42-48)
    std::vector<double> ac(g_.NZ), aw(g_.NZ);
    const double zd = hp.rayleigh_zd, zt = m.zt;
    auto ray = [&](double z) {
      if (zd < 0 || z <= zd || zt <= zd) return 0.0;
      const double sn = std::sin(1.5707963267948966 * (z - zd) / (zt - zd));
      return hp.rayleigh_alpha * sn * sn;
    };
    for (int k = 0; k < g_.NZ; ++k) {
      ac[k] = ray(m.zeta_c[k]);
      aw[k] = ray(m.zeta_w[k]);
    }
    met_.ray_c = upload_new<T>(ac.data(), s1);
    met_.ray_w = upload_new<T>(aw.data(), s1);
  }

  // Acoustic coefficients, formed in FP64 on the host exactly as
  // k_acou_coef_c / k_acou_coef_f (kernels.cu:452-474) do on the device.
  {
    std::vector<double> rtjc(hn), rtjx(hn), rtjy(hn), kt3(hn), aw(hn);
    for (int kr = 0; kr < g_.NZ; ++kr)
      for (int jr = 0; jr < g_.NY; ++jr)
        for (int ir = 0; ir < hNX_; ++ir) {
          const size_t c = ((size_t)kr * g_.NY + jr) * hNX_ + ir;
          const double J = m.jac[(size_t)jr * hNX_ + ir];
          rtjc[c] = b.rhob[c] * b.thvb[c] * J;
          kt3[c] = dev::kRd * b.pib[c] / (dev::kCv * b.rhob[c] * b.thvb[c] * J);
          aw[c] = b.rhobw[c] * b.thvbw[c];
        }
    for (int kr = 0; kr < g_.NZ; ++kr)
      for (int jr = 0; jr < g_.NY; ++jr)
        for (int ir = 0; ir < hNX_; ++ir) {
          const size_t c = ((size_t)kr * g_.NY + jr) * hNX_ + ir;
          rtjx[c] = ir > 0 ? 0.5 * (rtjc[c - 1] + rtjc[c]) : rtjc[c];
          rtjy[c] = jr > 0 ? 0.5 * (rtjc[c - hNX_] + rtjc[c]) : rtjc[c];
        }
    // Column-uniform profiles over every column of the padded grid at the
    // levels 0 .. nz-1 whose values the kernels use (equal bits, NaN
    // included).
    auto uniform = [&](const double* a) {
      if (!a) return true;
      for (int kr = gr.ng; kr < gr.ng + gr.nz; ++kr) {
        const double ref = a[((size_t)kr * g_.NY + gr.ng) * hNX_ + gr.ng];
        for (int jr = 0; jr < g_.NY; ++jr)
          for (int ir = 0; ir < hNX_; ++ir) {
            const double v = a[((size_t)kr * g_.NY + jr) * hNX_ + ir];
            if (std::memcmp(&v, &ref, sizeof v) != 0) return false;
          }
      }
      return true;
    };
    p_.hom = !p_.terrain && uniform(b.thb) && uniform(b.thbw) && uniform(b.dthbdz) &&
             uniform(b.qvb) && uniform(b.rhob) && uniform(b.rhobw) && uniform(b.pib) &&
             uniform(rtjx.data()) && uniform(rtjy.data()) && uniform(kt3.data()) &&
             uniform(aw.data());
    if (const char* e = std::getenv("WFE_UNIFORM_COLUMNS"))
      if (std::string(e) == "0") p_.hom = 0;
    base_.hom = p_.hom;
    coef_.rtjx = upload_new<T>(rtjx.data(), s3);
    coef_.rtjy = upload_new<T>(rtjy.data(), s3);
    coef_.kt3 = upload_new<T>(kt3.data(), s3);
    coef_.aw = upload_new<T>(aw.data(), s3);
  }

  tf_dev_ = arena_.alloc<T>(1);
  auto alloc_set = [&](StateSet<T>& s) {
    s.u = arena_.alloc<T>(n);
    s.v = arena_.alloc<T>(n);
    s.w = arena_.alloc<T>(n);
    s.thp = arena_.alloc<T>(n);
    s.qv = arena_.alloc<T>(n);
    s.qc = arena_.alloc<T>(n);
    s.qr = arena_.alloc<T>(n);
    s.pip = arena_.alloc<double>(n);
  };
  alloc_set(S_[0]);
  alloc_set(S_[1]);
  alloc_set(SE_[0]);
  alloc_set(SE_[1]);
  for (int q = 0; q < 2; ++q) {
    wu_[q] = arena_.alloc<T>(n);
    wv_[q] = arena_.alloc<T>(n);
    xthp_[q] = arena_.alloc<T>(n);
  }
  ww_ = arena_.alloc<T>(n);
  for (int q = 0; q < 3; ++q) xpip_[q] = arena_.alloc<double>(n);
  tu_ = arena_.alloc<T>(n);
  tv_ = arena_.alloc<T>(n);
  tw_ = arena_.alloc<T>(n);
  tth_ = arena_.alloc<T>(n);
  tpi_ = arena_.alloc<T>(n);
  if (p_.moisture) twf_ = arena_.alloc<T>(n);
  for (int q = 0; q < 3; ++q) tq_[q] = arena_.alloc<T>(n);
  mfx_ = arena_.alloc<T>(n);
  mfy_ = arena_.alloc<T>(n);
  mfz_ = arena_.alloc<T>(n);
  div_ = arena_.alloc<T>(n);
  if (p_.pd_moist)
    for (int q = 0; q < 3; ++q) {
      pratio_[q] = arena_.alloc<T>(n);
    }
  rain_ = arena_.alloc<T>(n2_);
  d_absmax_ = arena_.alloc<unsigned>(1);
  diag_nb_ = (int)(cdiv(g_.nx, 64) * cdiv(g_.ny, 4) * cdiv(g_.nz + 1, dev::kDiagLevels));
  diag_part_ = arena_.alloc<double>((size_t)(diag_nb_ + 1) * dev::kDgSlots);
  {
    int dev = 0;
    WFE_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop{};
    WFE_HIP(hipGetDeviceProperties(&prop, dev));
    build_methods(prop);
  }
  WFE_HIP(hipDeviceSynchronize());
}

template <class T>
T* Ctx<T>::field_ptr(int f) {
  StateSet<T>& s = S_[cur_];
  switch (f) {
    case WFE_F_U: return s.u;
    case WFE_F_V: return s.v;
    case WFE_F_W: return s.w;
    case WFE_F_THP: return s.thp;
    case WFE_F_QV: return s.qv;
    case WFE_F_QC: return s.qc;
    case WFE_F_QR: return s.qr;
    case WFE_F_TU: return tu_;
    case WFE_F_TV: return tv_;
    case WFE_F_TW: return tw_;
    case WFE_F_TTHP: return tth_;
    case WFE_F_TPIP: return tpi_;
    case WFE_F_TQV: return tq_[0];
    case WFE_F_TQC: return tq_[1];
    case WFE_F_TQR: return tq_[2];
    case WFE_F_RAIN: return rain_;
    default: throw ArgError("unknown field selector " + std::to_string(f));
  }
}

template <class T>
void Ctx<T>::upload(int f, const double* h, size_t n) {
  const bool two_d = f == WFE_F_RAIN;
  if (n != (two_d ? hn2_ : hn3_)) throw ArgError("upload: element count mismatch");
  WFE_HIP(hipStreamSynchronize(stream_));
  if (f == WFE_F_PIP) {
    const std::vector<double> tmp = to_dev<double>(h, true);
    WFE_HIP(hipMemcpy(S_[cur_].pip, tmp.data(), npts_ * sizeof(double), hipMemcpyHostToDevice));
    return;
  }
  T* d = field_ptr(f);
  const std::vector<T> tmp = to_dev<T>(h, !two_d);
  WFE_HIP(hipMemcpy(d, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
}

template <class T>
void Ctx<T>::download(int f, double* h, size_t n) {
  if (f >= WFE_F_TSK && f <= WFE_F_U10) {
    if (!phys_on_) throw ArgError("download: surface fields need the surface physics");
    if (n != hn2_) throw ArgError("download: element count mismatch");
    WFE_HIP(hipStreamSynchronize(stream_));
    std::vector<double> tmp(n2_);
    WFE_HIP(hipMemcpy(tmp.data(), phys_field(f), n2_ * sizeof(double), hipMemcpyDeviceToHost));
    to_host(tmp, h, false);
    return;
  }
  const bool two_d = f == WFE_F_RAIN;
  if (n != (two_d ? hn2_ : hn3_)) throw ArgError("download: element count mismatch");
  WFE_HIP(hipStreamSynchronize(stream_));
  if (f == WFE_F_PIP) {
    std::vector<double> tmp(npts_);
    WFE_HIP(hipMemcpy(tmp.data(), S_[cur_].pip, npts_ * sizeof(double), hipMemcpyDeviceToHost));
    to_host(tmp, h, true);
    return;
  }
  const T* d = field_ptr(f);
  std::vector<T> tmp(two_d ? n2_ : npts_);
  WFE_HIP(hipMemcpy(tmp.data(), d, tmp.size() * sizeof(T), hipMemcpyDeviceToHost));
  to_host(tmp, h, !two_d);
}

template <class T>
float Ctx<T>::absmax(int f) {
  if (f < 0 || f >= WFE_F_RAIN) throw ArgError("absmax: field must be a 3-D field");
  WFE_HIP(hipMemsetAsync(d_absmax_, 0, sizeof(unsigned), stream_));
  const int kmax = f == WFE_F_W ? g_.nz : g_.nz - 1;
  dim3 grd(cdiv(g_.nx, 64), cdiv(g_.ny, 4), (unsigned)(kmax + 1));
  if (f == WFE_F_PIP)
    dev::k_absmax<double><<<grd, dim3(64, 4), 0, stream_>>>(g_, S_[cur_].pip, kmax, d_absmax_);
  else
    dev::k_absmax<T><<<grd, dim3(64, 4), 0, stream_>>>(g_, field_ptr(f), kmax, d_absmax_);
  check_launch("absmax");
  unsigned bits = 0;
  WFE_HIP(hipMemcpyAsync(&bits, d_absmax_, sizeof(unsigned), hipMemcpyDeviceToHost, stream_));
  WFE_HIP(hipStreamSynchronize(stream_));
  float v;
  std::memcpy(&v, &bits, sizeof(float));
  if (std::isnan(v)) v = INFINITY;
  return v;
}

template <class T>
void Ctx<T>::diagnose(double* rec) {
  const StateSet<T>& s = S_[cur_];
  const dim3 grd(cdiv(g_.nx, 64), cdiv(g_.ny, 4), cdiv(g_.nz + 1, dev::kDiagLevels));
  dev::k_diag_partial<T><<<grd, dim3(64, 4), 0, stream_>>>(
      g_, s.w, s.thp, s.pip, p_.moisture ? s.qc : nullptr, p_.moisture ? s.qr : nullptr,
      diag_part_);
  check_launch("diag_partial");
  double* fin = diag_part_ + (size_t)diag_nb_ * dev::kDgSlots;
  dev::k_diag_final<<<1, 256, 0, stream_>>>(diag_part_, diag_nb_, fin);
  check_launch("diag_final");
  WFE_HIP(hipMemcpyAsync(rec, fin, dev::kDgSlots * sizeof(double), hipMemcpyDeviceToHost,
                         stream_));
  WFE_HIP(hipStreamSynchronize(stream_));
}

template <class T>
void Ctx<T>::pack_async(int f, float* dst, size_t n) {
  const bool is_state = f >= WFE_F_U && f <= WFE_F_QR;
  const bool is_sfc = f >= WFE_F_TSK && f <= WFE_F_U10;
  if (!is_state && f != WFE_F_RAIN && !is_sfc)
    throw ArgError("pack: field must be a state field, the rain or a surface field");
  if (is_sfc && !phys_on_) throw ArgError("pack: surface fields need the surface physics");
  if (!p_.moisture && (f == WFE_F_QV || f == WFE_F_QC || f == WFE_F_QR || f == WFE_F_RAIN))
    throw ArgError("pack: moisture fields of a dry context");
  const int nlev = f == WFE_F_W ? g_.nz + 1 : (f == WFE_F_RAIN || is_sfc ? 1 : g_.nz);
  if (n != (size_t)g_.nx * g_.ny * nlev) throw ArgError("pack: element count mismatch");
  if (!ostream_) {
    WFE_HIP(hipStreamCreate(&ostream_));
    for (int q = 0; q < WFE_F_COUNT; ++q) {
      WFE_HIP(hipEventCreateWithFlags(&out_packed_[q], hipEventDisableTiming));
      WFE_HIP(hipEventCreateWithFlags(&out_copied_[q], hipEventDisableTiming));
    }
  }
  if (!out_stage_[f]) {
    WFE_HIP(hipStreamSynchronize(stream_));
    out_stage_[f] = arena_.alloc<float>((size_t)g_.nx * g_.ny * (g_.nz + 1));
  }
  // Do not repack over a copy still in flight.
  if (out_used_[f]) WFE_HIP(hipStreamWaitEvent(stream_, out_copied_[f], 0));
  out_used_[f] = true;
  const dim3 grd(cdiv(g_.nx, 64), cdiv(g_.ny, 4), (unsigned)nlev);
  if (f == WFE_F_PIP)
    dev::k_pack_interior<double><<<grd, dim3(64, 4), 0, stream_>>>(g_, S_[cur_].pip, 0,
                                                                   out_stage_[f]);
  else if (is_sfc)
    dev::k_pack_interior<double><<<grd, dim3(64, 4), 0, stream_>>>(g_, phys_field(f), 1,
                                                                   out_stage_[f]);
  else
    dev::k_pack_interior<T><<<grd, dim3(64, 4), 0, stream_>>>(g_, field_ptr(f),
                                                              f == WFE_F_RAIN, out_stage_[f]);
  check_launch("pack_interior");
  WFE_HIP(hipEventRecord(out_packed_[f], stream_));
  WFE_HIP(hipStreamWaitEvent(ostream_, out_packed_[f], 0));
  WFE_HIP(hipMemcpyAsync(dst, out_stage_[f], n * sizeof(float), hipMemcpyDeviceToHost,
                         ostream_));
  WFE_HIP(hipEventRecord(out_copied_[f], ostream_));
}

template <class T>
long long Ctx<T>::output_mark() {
  std::lock_guard<std::mutex> lk(out_mu_);
  if (!ostream_) throw ArgError("output_mark: nothing has been packed");
  const long long t = out_seq_++;
  hipEvent_t& e = out_ring_[t % kOutRing];
  if (!e) WFE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  WFE_HIP(hipEventRecord(e, ostream_));
  return t;
}

template <class T>
void Ctx<T>::output_wait(long long t) {
  hipEvent_t e;
  {
    std::lock_guard<std::mutex> lk(out_mu_);
    if (t < 0 || t >= out_seq_ || t < out_seq_ - kOutRing)
      throw ArgError("output_wait: unknown or expired ticket");
    e = out_ring_[t % kOutRing];
  }
  WFE_HIP(hipEventSynchronize(e));
}

template <class T>
void Ctx<T>::set_bdy_weights(const double* w, size_t n) {
  if (n != hn2_) throw ArgError("bdy weights: element count mismatch");
  if (!bdy_wgt_) {
    bdy_wgt_ = arena_.alloc<T>(n2_);
    for (int v = 0; v < 5; ++v) {
      bdy_lo_[v] = arena_.alloc<T>(npts_);
      bdy_hi_[v] = arena_.alloc<T>(npts_);
    }
  }
  const std::vector<T> tmp = to_dev<T>(w, false);
  WFE_HIP(hipMemcpy(bdy_wgt_, tmp.data(), n2_ * sizeof(T), hipMemcpyHostToDevice));
  // Rim width: every interior column with a positive weight lies within
  // bdy_rim_ cells of an edge (0: relax over the whole grid).
  int rim = 0;
  for (int j = 0; j < g_.ny; ++j)
    for (int i = 0; i < g_.nx; ++i)
      if (w[(size_t)(j + g_.ng) * hNX_ + (i + g_.ng)] > 0.0)
        rim = std::max(rim, 1 + std::min(std::min(i, g_.nx - 1 - i), std::min(j, g_.ny - 1 - j)));
  bdy_rim_ = 2 * rim < std::min(g_.nx, g_.ny) ? rim : 0;
  bdy_active_ = true;
  ++gen_;
}

template <class T>
void Ctx<T>::set_bdy_target(int slot, int var, const double* h, size_t n) {
  if (!bdy_wgt_) throw ArgError("bdy target before bdy weights");
  if (slot < 0 || slot > 1 || var < 0 || var > 4 || n != hn3_)
    throw ArgError("bdy target: bad slot/var/size");
  WFE_HIP(hipStreamSynchronize(stream_));
  const std::vector<T> tmp = to_dev<T>(h, true);
  WFE_HIP(hipMemcpy(slot == 0 ? bdy_lo_[var] : bdy_hi_[var], tmp.data(), npts_ * sizeof(T),
                    hipMemcpyHostToDevice));
}

template <class T>
void Ctx<T>::fill_ghosts(const StateSet<T>& s) {
  dev::FillList<T> L{};
  dev::FillList<double> Lp{};
  auto add = [&](T* f, int kind, int xl, int yl) {
    L.f[L.n] = f;
    L.kind[L.n] = kind;
    L.xlast[L.n] = xl;
    L.ylast[L.n] = yl;
    ++L.n;
  };
  const int nx = g_.nx, ny = g_.ny;
  add(s.u, dev::kFillScalar, nx, ny - 1);
  add(s.v, dev::kFillScalar, nx - 1, ny);
  add(s.w, dev::kFillW, nx - 1, ny - 1);
  add(s.thp, dev::kFillScalar, nx - 1, ny - 1);
  if (p_.moisture) {
    add(s.qv, dev::kFillScalar, nx - 1, ny - 1);
    add(s.qc, dev::kFillScalar, nx - 1, ny - 1);
    add(s.qr, dev::kFillScalar, nx - 1, ny - 1);
  }
  if constexpr (std::is_same_v<T, double>) {
    add(s.pip, dev::kFillScalar, nx - 1, ny - 1);
  } else {
    Lp.f[0] = s.pip;
    Lp.kind[0] = dev::kFillScalar;
    Lp.xlast[0] = nx - 1;
    Lp.ylast[0] = ny - 1;
    Lp.n = 1;
  }
  const int ncell = dev::ghost_count(g_);
  TIMED("fill_ghosts", dev::k_fill_all<T><<<cdiv(ncell, 256), 256, 0, stream_>>>(g_, p_.bcx, p_.bcy, L, Lp));
  check_launch("fill_ghosts");
}

template <class T>
void Ctx<T>::slow_tendencies(const StateSet<T>& s, double dt, const StateSet<T>* qout,
                             double dt_stage, bool fold_buoyancy) {
  {
    // 64 x 4 whatever lv_bx_ (32-wide tiles measured slower for this kernel)
    const dim3 grd(cdiv(g_.nx + 5, 64), cdiv(g_.ny + 5, 4), (unsigned)g_.nz);
    TIMED("massflux", dev::k_massflux<T><<<grd, blk3(), 0, stream_>>>(g_, base_, met_, s.u, s.v, s.w, mfx_, mfy_,
                                                    mfz_, div_));
  }
  const int do_q = p_.moisture && !p_.pd_moist;
  const StateSet<T>& sn = S_[cur_];
  // Stage update of q where T_q is formed (qout), and the moist buoyancy of
  // the substeps (from the time-n moisture) added to T_w (fold_buoyancy).
  dev::QStage<T> qs{};
  if (qout) {
    qs.qn[0] = sn.qv;
    qs.qn[1] = sn.qc;
    qs.qn[2] = sn.qr;
    qs.qo[0] = qout->qv;
    qs.qo[1] = qout->qc;
    qs.qo[2] = qout->qr;
    qs.dts = (T)dt_stage;
    qs.on = 1;
  }
  const int stage_on = qout && do_q;
  const int buoy_on = fold_buoyancy && p_.moisture;
  T* tq0 = stage_on ? qout->qv : tq_[0];
  T* tq1 = stage_on ? qout->qc : tq_[1];
  T* tq2 = stage_on ? qout->qr : tq_[2];
  auto launch = [&](auto kz) {
    constexpr int KZ = decltype(kz)::value;
    const int ty = 256 / tend_tx_;
    const dim3 grd(cdiv(g_.nx, tend_tx_), cdiv(g_.ny, ty), (unsigned)cdiv(g_.nz, KZ));
    TIMED("tendencies",
          (dev::k_tendencies<T, KZ><<<grd, dim3(tend_tx_, ty), 0, stream_>>>(
              g_, p_, base_, met_, (T)dt, s.u, s.v, s.w, s.thp, s.qv, s.qc, s.qr, mfx_, mfy_, mfz_,
              div_, tu_, tv_, tw_, tth_, tq0, tq1, tq2, sn.qv, sn.qc, sn.qr, do_q, (T)dt_stage,
              stage_on, buoy_on)));
  };
  auto launch_cur = [&] {
    if (tend_levels_ == 4)
      launch(std::integral_constant<int, 4>{});
    else if (tend_levels_ == 2)
      launch(std::integral_constant<int, 2>{});
    else
      launch(std::integral_constant<int, 1>{});
  };
  if (!tend_tuned_) {
    // The launch is a pure function of the stage's inputs, so repeating it
    // for timing leaves the step's results unchanged.
    tend_report_.clear();
    float best = 1e30f;
    int best_l = tend_levels_, best_t = tend_tx_;
    for (const int lv : {1, 2, 4})
      for (const int tx : {64, 32}) {
        tend_levels_ = lv;
        tend_tx_ = tx;
        const float ms = time_launches(launch_cur);
        char buf[96];
        std::snprintf(buf, sizeof(buf), "tcand:%dx%d 1 %.6f\n", lv, tx, ms);
        tend_report_ += buf;
        if (ms < best) {
          best = ms;
          best_l = lv;
          best_t = tx;
        }
      }
    tend_levels_ = best_l;
    tend_tx_ = best_t;
    tend_report_ += "ttuned:" + std::to_string(best_l) + "x" + std::to_string(best_t) + " 0 0\n";
    tend_tuned_ = true;
    ++gen_;
  }
  launch_cur();
  if (p_.moisture && p_.pd_moist) {
    dev::PDSet<T> S{};
    const T* q[3] = {s.qv, s.qc, s.qr};
    for (int q3 = 0; q3 < 3; ++q3) {
      S.q[q3] = q[q3];
      S.ratio[q3] = pratio_[q3];
      S.tq[q3] = tq_[q3];
    }
    const dim3 gi = lv_grid(g_.nx, g_.ny, g_.nz);
    TIMED("pd_ratio", dev::k_pd_ratio<T><<<gi, lv_block(), 0, stream_>>>(
                          g_, base_, met_, (T)dt, mfx_, mfy_, mfz_, S));
    TIMED("pd_apply", dev::k_pd_apply<T><<<gi, lv_block(), 0, stream_>>>(
                          g_, p_, base_, met_, mfx_, mfy_, mfz_, div_, S, qs));
  }
  check_launch("slow_tendencies");
}

template <class T>
void Ctx<T>::bdy_relax(const StateSet<T>& s) {
  dev::RelaxSet<T> R{};
  R.f[0] = s.u;
  R.f[1] = s.v;
  R.f[2] = s.thp;
  R.f[3] = s.qv;
  R.pip = s.pip;
  for (int v = 0; v < 5; ++v) {
    R.lo[v] = bdy_lo_[v];
    R.hi[v] = bdy_hi_[v];
  }
  R.t[0] = tu_;
  R.t[1] = tv_;
  R.t[2] = tth_;
  R.t[3] = tpi_;
  R.t[4] = tq_[0];
  R.wgt = bdy_wgt_;
  R.tf = tf_dev_;
  if (bdy_rim_ > 0) {
    const int ncol = 2 * bdy_rim_ * g_.nx + 2 * bdy_rim_ * (g_.ny - 2 * bdy_rim_);
    TIMED("bdy_relax", dev::k_bdy_relax_rim<T><<<dim3(cdiv(ncol, 256), (unsigned)g_.nz), 256, 0,
                                                 stream_>>>(g_, R, bdy_rim_));
  } else {
    const dim3 grd = lv_grid(g_.nx, g_.ny, g_.nz);
    TIMED("bdy_relax", dev::k_bdy_relax<T><<<grd, lv_block(), 0, stream_>>>(g_, R));
  }
  check_launch("bdy_relax");
}

// One RK3 stage of the fast system from the time-n state S_[cur_] into out.
// Column-solve arguments of a split substep from the common substep args.
template <class T>
dev::WPArgs<T> split_wp_impl(const dev::FusedArgs<T>& a) {
  dev::WPArgs<T> w{};
  w.u = a.uv.u_dst;
  w.v = a.uv.v_dst;
  w.w_src = a.w_src;
  w.pip_src = a.uv.pcur;
  w.thp_src = a.uv.thp_cur;
  w.w_dst = a.w_dst;
  w.pip_dst = a.pip_dst;
  w.thp_dst = a.thp_dst;
  w.tw = a.tw;
  w.tth = a.tth;
  w.tpi = a.tpi;
  w.tpw = a.tpw;
  w.qv = a.uv.qv;
  w.dtau = (double)a.uv.dtau;
  return w;
}

template <class T>
void Ctx<T>::batch_member(dev::BatchMember<T>& M, double dt, int m, int s) const {
  const int nsub = stage_nsub(m);
  const dev::FusedArgs<T> a = substep_args(stage_out(m), stage_dt(dt, m) / nsub, s, nsub);
  M.g = g_;
  M.p = p_;
  M.b = base_;
  M.m = met_;
  M.cf = coef_;
  M.uv = a.uv;
  M.wp = split_wp_impl<T>(a);
}

template <class T>
template <int TX, int NW, int KPT, int WPE>
void Ctx<T>::add_split_tile(const char* name, size_t cap) {
  using Tile = dev::WpiTile<TX, NW, KPT>;
  if (g_.nz > NW * KPT) return;
  const size_t lb = Tile::lds_bytes(g_.nz);
  // Instantiations by (moisture, pi' tendency present).
  const void* fns[4] = {reinterpret_cast<const void*>(&dev::k_acou_wpi_tile<T, TX, NW, KPT, false, false, WPE>),
                        reinterpret_cast<const void*>(&dev::k_acou_wpi_tile<T, TX, NW, KPT, false, true, WPE>),
                        reinterpret_cast<const void*>(&dev::k_acou_wpi_tile<T, TX, NW, KPT, true, false, WPE>),
                        reinterpret_cast<const void*>(&dev::k_acou_wpi_tile<T, TX, NW, KPT, true, true, WPE>)};
  if (lb > 65536) {
    if (lb > cap) return;
    for (const void* f : fns)
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb) != hipSuccess) {
        (void)hipGetLastError();
        return;
      }
  }
  methods_.push_back({name, [this, lb](const dev::FusedArgs<T>& a) {
                        launch_uv_tile(a.uv);
                        const dev::WPArgs<T> w = split_wp_impl<T>(a);
                        const dim3 gt(cdiv(g_.nx * g_.ny, TX));
                        const dim3 bt(TX, NW);
                        const bool moist = p_.moisture, tpi = w.tpi != nullptr;
                        if (moist && tpi)
                          TIMED("acou_wpi", (dev::k_acou_wpi_tile<T, TX, NW, KPT, true, true, WPE>
                                             <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, w)));
                        else if (moist)
                          TIMED("acou_wpi", (dev::k_acou_wpi_tile<T, TX, NW, KPT, true, false, WPE>
                                             <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, w)));
                        else if (tpi)
                          TIMED("acou_wpi", (dev::k_acou_wpi_tile<T, TX, NW, KPT, false, true, WPE>
                                             <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, w)));
                        else
                          TIMED("acou_wpi", (dev::k_acou_wpi_tile<T, TX, NW, KPT, false, false, WPE>
                                             <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, w)));
                      }});
}

template <class T>
template <int TX, int NW, int KPT, int WPE>
void Ctx<T>::add_row(const char* name, size_t cap) {
  using Shape = dev::RowShape<TX, NW, KPT>;
  if (g_.nz > NW * KPT) return;
  const size_t lb = Shape::lds_bytes(g_.nz, p_.terrain != 0, sizeof(T));
  // Instantiations by (moisture, pi' tendency present).
  const void* fns[4] = {reinterpret_cast<const void*>(&dev::k_acou_row<T, TX, NW, KPT, false, false, WPE>),
                        reinterpret_cast<const void*>(&dev::k_acou_row<T, TX, NW, KPT, false, true, WPE>),
                        reinterpret_cast<const void*>(&dev::k_acou_row<T, TX, NW, KPT, true, false, WPE>),
                        reinterpret_cast<const void*>(&dev::k_acou_row<T, TX, NW, KPT, true, true, WPE>)};
  if (lb > 65536) {
    if (lb > cap) return;
    for (const void* f : fns)
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb) != hipSuccess) {
        (void)hipGetLastError();
        return;
      }
  }
  methods_.push_back({name, [this, lb](const dev::FusedArgs<T>& a) {
                        const dim3 gt(cdiv(g_.nx, Shape::NC), (unsigned)g_.ny);
                        const dim3 bt(TX, NW);
                        const bool moist = p_.moisture, tpi = a.tpi != nullptr;
                        if (moist && tpi)
                          TIMED("acoustic_fused", (dev::k_acou_row<T, TX, NW, KPT, true, true, WPE>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a)));
                        else if (moist)
                          TIMED("acoustic_fused", (dev::k_acou_row<T, TX, NW, KPT, true, false, WPE>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a)));
                        else if (tpi)
                          TIMED("acoustic_fused", (dev::k_acou_row<T, TX, NW, KPT, false, true, WPE>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a)));
                        else
                          TIMED("acoustic_fused", (dev::k_acou_row<T, TX, NW, KPT, false, false, WPE>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a)));
                      }});
}

template <class T>
template <int CX, int CY, int NW, int KPT>
void Ctx<T>::add_tile2d(const char* name, size_t cap) {
  using Sh = dev::Tile2D<CX, CY>;
  if (g_.nz > NW * KPT) return;
  const size_t lb = Sh::template lds_bytes<T>(g_.nz);
  const void* fns[4] = {reinterpret_cast<const void*>(&dev::k_acou_tile2d<T, CX, CY, NW, KPT, false, false>),
                        reinterpret_cast<const void*>(&dev::k_acou_tile2d<T, CX, CY, NW, KPT, false, true>),
                        reinterpret_cast<const void*>(&dev::k_acou_tile2d<T, CX, CY, NW, KPT, true, false>),
                        reinterpret_cast<const void*>(&dev::k_acou_tile2d<T, CX, CY, NW, KPT, true, true>)};
  if (lb > 65536) {
    if (lb > cap) return;
    for (const void* f : fns)
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb) != hipSuccess) {
        (void)hipGetLastError();
        return;
      }
  }
  methods_.push_back({name, [this, lb](const dev::FusedArgs<T>& a) {
                        const dev::WPArgs<T> w = split_wp_impl<T>(a);
                        const dim3 gt(cdiv(g_.nx, CX), cdiv(g_.ny, CY));
                        const dim3 bt(64, NW);
                        const bool moist = p_.moisture, tpi = a.tpi != nullptr;
                        if (moist && tpi)
                          TIMED("acoustic_fused", (dev::k_acou_tile2d<T, CX, CY, NW, KPT, true, true>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a.uv, w)));
                        else if (moist)
                          TIMED("acoustic_fused", (dev::k_acou_tile2d<T, CX, CY, NW, KPT, true, false>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a.uv, w)));
                        else if (tpi)
                          TIMED("acoustic_fused", (dev::k_acou_tile2d<T, CX, CY, NW, KPT, false, true>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a.uv, w)));
                        else
                          TIMED("acoustic_fused", (dev::k_acou_tile2d<T, CX, CY, NW, KPT, false, false>
                                                   <<<gt, bt, lb, stream_>>>(g_, p_, base_, met_, coef_, a.uv, w)));
                      }});
}

template <class T>
void Ctx<T>::build_methods(const hipDeviceProp_t& prop) {
  const size_t cap = (size_t)prop.sharedMemPerBlock;
  // General split pair: one thread per column (any nz <= 319).
  methods_.push_back({"split-col", [this](const dev::FusedArgs<T>& a) {
                        launch_uv_tile(a.uv);
                        const dev::WPArgs<T> w = split_wp_impl<T>(a);
                        const dim3 gcol(cdiv(g_.nx, 64), (unsigned)g_.ny);
                        if (g_.nz + 1 <= 64)
                          TIMED("acou_wpi", dev::k_acou_wpi_col<T, 64>
                                            <<<gcol, dim3(64, 1), 0, stream_>>>(g_, p_, base_,
                                                                                met_, coef_, w));
                        else
                          TIMED("acou_wpi", dev::k_acou_wpi_col<T, 320>
                                            <<<gcol, dim3(64, 1), 0, stream_>>>(g_, p_, base_,
                                                                                met_, coef_, w));
                      }});
  add_split_tile<32, 8, 6>("split-32x8", cap);
  add_split_tile<32, 16, 3>("split-32x16", cap);
  add_split_tile<64, 8, 6>("split-64x8", cap);
  add_split_tile<64, 16, 3>("split-64x16", cap);
  add_split_tile<16, 16, 3>("split-16x16", cap);
  add_split_tile<32, 16, 4>("split-32x16x4", cap);  // columns up to 64 levels
  add_split_tile<32, 16, 3, 4>("split-32x16w4", cap);   // two blocks per CU
  add_split_tile<16, 16, 3, 4>("split-16x16w4", cap);
  add_row<64, 16, 3>("fused-64x16", cap);
  add_row<64, 14, 3>("fused-64x14", cap);     // 42 level slots (nz <= 42)
  add_row<64, 10, 4>("fused-64x10x4", cap);   // 40 level slots (nz <= 40)
  add_row<64, 8, 6>("fused-64x8", cap);
  add_row<32, 16, 3>("fused-32x16", cap);
  add_row<32, 8, 6>("fused-32x8", cap);
  add_row<32, 16, 4>("fused-32x16x4", cap);  // columns up to 64 levels
  add_row<32, 16, 3, 4>("fused-32x16w4", cap);  // two blocks per CU
  // 2-D column tiles (terrain): 16 x 4 and 8 x 8 columns per workgroup
  add_tile2d<16, 4, 16, 3>("tile2d-16x4", cap);
  add_tile2d<8, 8, 16, 3>("tile2d-8x8", cap);
  // Selection: WFE_ACOUSTIC_KERNEL = exact method name, or a family prefix
  // ("split", "fused"); unset or "auto" = all. WFE_WPI_KERNEL=column keeps
  // only the one-thread-per-column pair.
  std::string sel;
  if (const char* e = std::getenv("WFE_ACOUSTIC_KERNEL")) sel = e;
  if (const char* e = std::getenv("WFE_WPI_KERNEL"))
    if (std::string(e) == "column") sel = "split-col";
  if (!sel.empty() && sel != "auto") {
    std::vector<SubstepMethod> keep;
    for (auto& mth : methods_)
      if (mth.name == sel || mth.name.rfind(sel + "-", 0) == 0) keep.push_back(mth);
    if (keep.empty()) throw ArgError("WFE_ACOUSTIC_KERNEL=" + sel + " matches no method for this grid");
    methods_.swap(keep);
  }
  if (methods_.size() == 1) {
    method_ = 0;
    tuned_ = true;
    ++gen_;
  }
}

// Times every candidate substep method on the current state (writing only
// work buffers) and keeps the fastest; reported as "tuned:<name>" and
// "cand:<name> 1 <ms>" lines by wfe_kernel_timing_report.
template <class T>
void Ctx<T>::tune_acoustic(double dtau) {
  const StateSet<T>& sn = S_[cur_];
  dev::FusedArgs<T> a{};
  a.uv.u_src = sn.u;
  a.uv.v_src = sn.v;
  a.uv.u_dst = wu_[0];
  a.uv.v_dst = wv_[0];
  a.uv.pcur = sn.pip;
  a.uv.pprev = sn.pip;
  a.uv.pn = sn.pip;
  a.uv.thp_cur = sn.thp;
  a.uv.thp_n = sn.thp;
  a.uv.qv = sn.qv;
  a.uv.tu = tu_;
  a.uv.tv = tv_;
  a.uv.dtau = (T)dtau;
  a.w_src = sn.w;
  a.w_dst = ww_;
  a.pip_dst = xpip_[0];
  a.thp_dst = xthp_[0];
  a.tw = tw_;
  a.tth = tth_;
  a.tpi = nullptr;
  hipEvent_t e0, e1;
  WFE_HIP(hipEventCreate(&e0));
  WFE_HIP(hipEventCreate(&e1));
  const bool was_on = timer_.on;
  timer_.on = false;
  float best = 1e30f;
  int best_m = 0;
  tune_report_.clear();
  for (int q = 0; q < (int)methods_.size(); ++q) {
    for (int r = 0; r < 2; ++r) methods_[q].run(a);
    WFE_HIP(hipEventRecord(e0, stream_));
    for (int r = 0; r < 5; ++r) methods_[q].run(a);
    WFE_HIP(hipEventRecord(e1, stream_));
    WFE_HIP(hipEventSynchronize(e1));
    check_launch("tune_acoustic");
    float ms = 0;
    WFE_HIP(hipEventElapsedTime(&ms, e0, e1));
    char buf[128];
    std::snprintf(buf, sizeof(buf), "cand:%s 1 %.6f\n", methods_[q].name.c_str(), ms / 5);
    tune_report_ += buf;
    if (ms < best) {
      best = ms;
      best_m = q;
    }
  }
  timer_.on = was_on;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  method_ = best_m;
  tuned_ = true;
  ++gen_;
}

// One RK3 stage of the fast system from the time-n state S_[cur_] into out.
// pi' rotates over three work buffers: the input of substep s is the output
// of s-1 and pi'prev is the input of s-1 (time n for s <= 1); theta' and
// u, v ping-pong; w is updated in place (own column only).
template <class T>
void Ctx<T>::acoustic_stage(const StateSet<T>& out, double dtau, int nsub) {
  if (!tuned_) tune_acoustic(dtau);
  for (int s = 0; s < nsub; ++s) methods_[method_].run(substep_args(out, dtau, s, nsub));
  check_launch("acoustic_stage");
}

// The buffers of substep s of nsub: u, v and theta' ping-pong, pi' rotates
// over three buffers so that the substep's pi*(pi', pi'prev) inputs are never
// its output; the last substep writes the stage output.
template <class T>
dev::FusedArgs<T> Ctx<T>::substep_args(const StateSet<T>& out, double dtau, int s,
                                     int nsub) const {
  const StateSet<T>& sn = S_[cur_];
  const bool first = s == 0, last = s == nsub - 1;
  dev::FusedArgs<T> a{};
  a.uv.u_src = first ? sn.u : wu_[(s - 1) & 1];
  a.uv.v_src = first ? sn.v : wv_[(s - 1) & 1];
  a.uv.u_dst = last ? out.u : wu_[s & 1];
  a.uv.v_dst = last ? out.v : wv_[s & 1];
  a.uv.pcur = first ? sn.pip : xpip_[(s - 1) % 3];
  a.uv.pprev = s <= 1 ? sn.pip : xpip_[(s - 2) % 3];
  a.uv.pn = sn.pip;
  a.uv.thp_cur = first ? sn.thp : xthp_[(s - 1) & 1];
  a.uv.thp_n = sn.thp;
  a.uv.qv = sn.qv;
  a.uv.tu = tu_;
  a.uv.tv = tv_;
  a.uv.dtau = (T)dtau;
  a.w_src = first ? sn.w : ww_;
  a.w_dst = last ? out.w : ww_;
  a.pip_dst = last ? out.pip : xpip_[s % 3];
  a.thp_dst = last ? out.thp : xthp_[s & 1];
  a.tw = tw_sub_ ? tw_sub_ : tw_;
  a.tth = tth_;
  a.tpi = bdy_active_ ? tpi_ : nullptr;
  a.tpw = bdy_wgt_;
  return a;
}

template <class T>
void Ctx<T>::stage_moisture(const StateSet<T>& out, double dt_stage) {
  const StateSet<T>& sn = S_[cur_];
  const dim3 grd = lv_grid(g_.nx, g_.ny, g_.nz);
  TIMED("stage_q", dev::k_stage_q<T><<<grd, lv_block(), 0, stream_>>>(g_, (T)dt_stage, sn.qv, sn.qc, sn.qr,
                                                       tq_[0], tq_[1], tq_[2], out.qv, out.qc,
                                                       out.qr));
  check_launch("stage_moisture");
}

template <class T>
void Ctx<T>::kessler_step(const StateSet<T>& s, double dt) {
  dev::KesArgs<T> a{s.thp, s.pip, s.qv, s.qc, s.qr, rain_, (T)dt};  // the tuner swaps it
  const dim3 gcol(cdiv(g_.nx, 64), (unsigned)g_.ny);
  // Tile shapes: TX columns, NW thread rows, KPT = 48 / NW levels per thread.
  auto tile = [&](auto tx, auto nw) {
    constexpr int TX = decltype(tx)::value, NW = decltype(nw)::value;
    const dim3 gt(cdiv(g_.nx, TX), (unsigned)g_.ny);
    TIMED("kessler", (dev::k_kessler_tile<T, TX, NW, 48 / NW><<<gt, dim3(TX, NW),
                                                                 dev::kessler_tile_lds<TX, NW, T>(g_.nz),
                                                                 stream_>>>(g_, base_, met_, a)));
  };
  using I8 = std::integral_constant<int, 8>;
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  auto run_tile = [&] {
    if (kes_tile_ == "32x8")
      tile(I32{}, I8{});
    else if (kes_tile_ == "32x16")
      tile(I32{}, I16{});
    else if (kes_tile_ == "16x16")
      tile(I16{}, I16{});
    else if (kes_tile_ == "64x16")
      tile(I64{}, I16{});
    else if (kes_tile_ == "32x24")
      tile(I32{}, std::integral_constant<int, 24>{});
    else if (kes_tile_ == "16x48")
      tile(I16{}, std::integral_constant<int, 48>{});
    else if (kes_tile_ == "64x8")
      tile(I64{}, I8{});
    else
      tile(I32{}, I16{});
  };
  if (g_.nz <= 48) {
    if (!kes_tuned_) {
      // Timed on a copy of the state (the stage estimates are free after
      // the three stages) with a scratch rain field, then run on the state.
      const StateSet<T>& sc = SE_[0];
      for (const auto& f : {std::make_pair(sc.thp, s.thp), std::make_pair(sc.qv, s.qv),
                            std::make_pair(sc.qc, s.qc), std::make_pair(sc.qr, s.qr)})
        WFE_HIP(hipMemcpyAsync(f.first, f.second, npts_ * sizeof(T), hipMemcpyDeviceToDevice,
                               stream_));
      if (!kes_rain_scratch_) kes_rain_scratch_ = arena_.alloc<T>(n2_);
      const dev::KesArgs<T> real = a;
      a = dev::KesArgs<T>{sc.thp, s.pip, sc.qv, sc.qc, sc.qr, kes_rain_scratch_, (T)dt};
      kes_report_.clear();
      float best = 1e30f;
      std::string best_t = kes_tile_;
      for (const char* t : {"32x16", "16x16", "32x8", "64x8", "64x16", "32x24", "16x48"}) {
        kes_tile_ = t;
        const float ms = time_launches(run_tile);
        kes_report_ += std::string("kcand:") + t + " 1 " + std::to_string(ms) + "\n";
        if (ms < best) {
          best = ms;
          best_t = t;
        }
      }
      kes_tile_ = best_t;
      kes_report_ += "ktuned:" + best_t + " 0 0\n";
      kes_tuned_ = true;
      ++gen_;
      a = real;
    }
    run_tile();
  } else if (g_.nz + 1 <= 64)
    TIMED("kessler", dev::k_kessler_col<T, 64><<<gcol, dim3(64, 1), 0, stream_>>>(g_, base_, met_, a));
  else
    TIMED("kessler", dev::k_kessler_col<T, 320><<<gcol, dim3(64, 1), 0, stream_>>>(g_, base_, met_, a));
  check_launch("kessler");
}

// Average launch time of one kernel of the step, launched back to back
// with the arguments it has in the first acoustic substep / the first stage
// of a step. Every kernel it can time writes only work buffers, tendencies,
// fluxes or ghost cells that the next step recomputes before reading them.
template <class T>
double Ctx<T>::kernel_bench(const std::string& name, int reps) {
  if (reps < 1) throw ArgError("kernel_bench: reps must be positive");
  WFE_HIP(hipStreamSynchronize(stream_));
  const StateSet<T>& sn = S_[cur_];
  solo_ = SoloBench{};
  solo_.name = name.c_str();
  solo_.reps = reps;
  const bool timer_was_on = timer_.on;
  timer_.on = false;
  if (name == "acou_uv" || name == "acou_wpi" || name == "acoustic_fused") {
    if (!tuned_) tune_acoustic(1.0);
    dev::FusedArgs<T> a{};
    a.uv.u_src = sn.u;
    a.uv.v_src = sn.v;
    a.uv.u_dst = wu_[0];
    a.uv.v_dst = wv_[0];
    a.uv.pcur = sn.pip;
    a.uv.pprev = sn.pip;
    a.uv.pn = sn.pip;
    a.uv.thp_cur = sn.thp;
    a.uv.thp_n = sn.thp;
    a.uv.qv = sn.qv;
    a.uv.tu = tu_;
    a.uv.tv = tv_;
    a.uv.dtau = (T)1.0;
    a.w_src = sn.w;
    a.w_dst = ww_;
    a.pip_dst = xpip_[0];
    a.thp_dst = xthp_[0];
    a.tw = tw_;
    a.tth = tth_;
    a.tpi = bdy_active_ ? tpi_ : nullptr;
    a.tpw = bdy_wgt_;
    methods_[method_].run(a);
  } else if (name == "fill_ghosts") {
    fill_ghosts(sn);
  } else if (name == "massflux" || name == "tendencies") {
    slow_tendencies(sn, 1.0);
  } else if (name == "stage_q") {
    stage_moisture(SE_[0], 1.0);
  }
  WFE_HIP(hipStreamSynchronize(stream_));
  timer_.on = timer_was_on;
  const double us = solo_.avg_us;
  solo_ = SoloBench{};
  if (us < 0) throw ArgError("kernel_bench: '" + name + "' is not launched by this context");
  return us;
}

// Integrator::step (integrator.cpp:71-111) with fixed-role buffers: the
// time-n set S_[cur_], the stage estimate SE_, the next time level
// S_[cur_^1], and ping-pong work buffers inside the acoustic loop.
template <class T>
void Ctx<T>::set_clock(double dt, double t) {
  if (!phys_on_) return;
  // hour of day at the end of the step (surface.cu:501-502)
  double utc = phys_start_hour_ + (t + dt) / 3600;
  utc = utc - 24 * std::floor(utc / 24);
  dev::k_set_scalar<double><<<1, 64, 0, stream_>>>(utc_dev_, utc);
  check_launch("set_utc");
}

template <class T>
void Ctx<T>::step(double dt, double t) {
  set_clock(dt, t);
  // The first step tunes the substep method and the tendency and Kessler
  // shapes (synchronous timing), so it and any step under per-kernel timing
  // run eagerly.
  if (!use_graph_ || !ready_for_capture()) {
    step_launches(dt);
    return;
  }
  const int par = cur_;
  if (!gexec_[par] || gdt_[par] != dt || ggen_[par] != gen_) {
    if (gexec_[par]) {
      WFE_HIP(hipGraphExecDestroy(gexec_[par]));
      gexec_[par] = nullptr;
    }
    hipGraph_t gr = nullptr;
    WFE_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      const CaptureScope cap;
      step_launches(dt);
    } catch (...) {
      (void)hipStreamEndCapture(stream_, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      cur_ = par;
      throw;
    }
    cur_ = par;
    WFE_HIP(hipStreamEndCapture(stream_, &gr));
    const hipError_t e = hipGraphInstantiate(&gexec_[par], gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      gexec_[par] = nullptr;
      throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    gdt_[par] = dt;
    ggen_[par] = gen_;
  }
  WFE_HIP(hipGraphLaunch(gexec_[par], stream_));
  cur_ ^= 1;
}

template <class T>
void Ctx<T>::step_launches(double dt) {
  for (int m = 0; m < 3; ++m) {
    stage_pre(dt, m);
    acoustic_stage(stage_out(m), stage_dt(dt, m) / stage_nsub(m), stage_nsub(m));
    stage_post(dt, m);
  }
  cur_ ^= 1;
  step_post(dt);
}

template <class T>
void Ctx<T>::surface_step(const StateSet<T>& s, double dt) {
  dev::SfcArgs a = sfc_;
  a.dt = dt;
  const dev::SfcState<T> st{s.thp, s.qv, s.u, s.v, s.qc, s.pip};
  auto run = [&](auto kmax) {
    constexpr int KMAX = decltype(kmax)::value;
    TIMED("sfc_scalar", (dev::k_sfc_scalar<T, KMAX><<<dim3(cdiv(g_.nx, 64), (unsigned)g_.ny), 64,
                                                       0, stream_>>>(g_, base_, met_, st, a)));
    if (g_.nx > 1)
      TIMED("sfc_u", (dev::k_sfc_face<T, KMAX, 0><<<dim3(cdiv(g_.nx - 1, 64), (unsigned)g_.ny), 64,
                                                    0, stream_>>>(g_, base_, met_, s.u, a)));
    if (g_.ny > 1)
      TIMED("sfc_v", (dev::k_sfc_face<T, KMAX, 1><<<dim3(cdiv(g_.nx, 64), (unsigned)(g_.ny - 1)),
                                                    64, 0, stream_>>>(g_, base_, met_, s.v, a)));
  };
  if (g_.nz <= 64)
    run(std::integral_constant<int, 64>{});
  else
    run(std::integral_constant<int, 320>{});
  check_launch("surface_step");
}

template <class T>
double* Ctx<T>::phys_field(int f) const {
  switch (f) {
    case WFE_F_TSK: return sfc_.tsk;
    case WFE_F_PBLH: return sfc_.pblh;
    case WFE_F_T2M: return sfc_.t2m;
    case WFE_F_U10: return sfc_.u10;
    default: return nullptr;
  }
}

// Integrator::set_physics + SfcPBL::init (surface.cu:440-482): the surface
// fields of the input, deep soil and all four soil layers at the initial
// surface temperature.
template <class T>
void Ctx<T>::set_physics(const wfe_sfc& sf) {
  const double* need[] = {sf.tsk, sf.land, sf.lat, sf.lon, sf.soilw};
  for (const double* q : need)
    if (!q) throw ArgError("set_physics: missing surface field");
  if (sf.doy < 1 || sf.doy > 366) throw ArgError("set_physics: day of year out of range");
  WFE_HIP(hipStreamSynchronize(stream_));
  if (!phys_on_) {
    for (double** f : {&sfc_.tsk, &sfc_.tdeep, &sfc_.land, &sfc_.lat, &sfc_.lon, &sfc_.soilw,
                       &sfc_.cdv, &sfc_.pblh, &sfc_.t2m, &sfc_.u10})
      *f = arena_.alloc<double>(n2_);
    sfc_.km = arena_.alloc<double>(npts_);
    sfc_.soilt = arena_.alloc<double>(4 * n2_);
    utc_dev_ = arena_.alloc<double>(1);
    sfc_.utc = utc_dev_;
  }
  auto up = [&](double* d, const double* h) {
    const std::vector<double> tmp = to_dev<double>(h, false);
    WFE_HIP(hipMemcpy(d, tmp.data(), n2_ * sizeof(double), hipMemcpyHostToDevice));
  };
  up(sfc_.tsk, sf.tsk);
  up(sfc_.tdeep, sf.tsk);
  up(sfc_.land, sf.land);
  up(sfc_.lat, sf.lat);
  up(sfc_.lon, sf.lon);
  up(sfc_.soilw, sf.soilw);
  for (int L = 0; L < 4; ++L) up(sfc_.soilt + (size_t)L * n2_, sf.tsk);
  sfc_.doy = sf.doy;
  sfc_.nonlocal = sf.nonlocal ? 1 : 0;
  phys_start_hour_ = sf.start_hour_utc;
  phys_on_ = true;
  ++gen_;
}

template <class T>
void Ctx<T>::acoustic_substeps(double dtau, int nsub) {
  const StateSet<T>& out = S_[cur_ ^ 1];
  if (p_.moisture) {  // the stand-alone T_w lacks the moist buoyancy
    const StateSet<T>& sn = S_[cur_];
    dev::MoistBuoy<T> mb{sn.qv, sn.qc, sn.qr, 1};
    if (g_.nz > 1) {
      const dim3 grd(cdiv(g_.nx, 64), cdiv(g_.ny, 4), (unsigned)(g_.nz - 1));
      dev::k_fold_moist_buoyancy<T><<<grd, blk3(), 0, stream_>>>(g_, base_, mb, tw_, twf_);
      check_launch("fold_moist_buoyancy");
    }
    tw_sub_ = twf_;
  }
  acoustic_stage(out, dtau, nsub);
  tw_sub_ = nullptr;
  const StateSet<T>& sn = S_[cur_];
  WFE_HIP(hipMemcpyAsync(out.qv, sn.qv, npts_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
  WFE_HIP(hipMemcpyAsync(out.qc, sn.qc, npts_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
  WFE_HIP(hipMemcpyAsync(out.qr, sn.qr, npts_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
  cur_ ^= 1;
}

// ------------------------------------------------------------------ adjoint
// Tangent-linear and adjoint of the dry large step: the forward step is
// re-run with every stage estimate, tendency, mass flux and substep output
// kept (the trajectory), then the TL kernels run forward or the AD kernels
// backward through the same sequence. Gradients over many steps keep the
// state every `ckpt` steps and recompute each segment.
//
// The trajectory is held in the context's own storage type T (FP32 fields
// and FP64 pi' in a mixed context, FP64 throughout in an FP64 one), so the
// linearisation is about the trajectory the context itself computes; the
// perturbations, adjoint variables, their accumulations and the objective
// are FP64 in both.
class AdjointBase {
 public:
  virtual ~AdjointBase() = default;
  // Host arrays: nf = 5 (u, v, w, theta', pi') or 6 (and q_v); the q_v
  // entry is ignored on input and zero on output when moisture is off.
  virtual void tl_step_host(double dt, const double* const* dx, double* const* dy, int nf) = 0;
  virtual void ad_step_host(double dt, const double* const* yb, double* const* xb, int nf) = 0;
  virtual double gradient(double dt, int nsteps, int ckpt, const double* wgt2d,
                          double* const* grad, int nf) = 0;
};

template <class T>
class Adjoint final : public AdjointBase {
 public:
  // Fields u, v, w, theta', pi' in storage type U (pi' always FP64), and
  // q_v when moisture is on (null otherwise).
  template <class U>
  struct Set5T {
    U *u = nullptr, *v = nullptr, *w = nullptr, *th = nullptr;
    double* pip = nullptr;
    U* q = nullptr;
  };
  using Set5 = Set5T<double>;  // perturbations / adjoint variables
  using TSet = Set5T<T>;       // trajectory / model states
  template <class U>
  struct Set4T {
    U *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  };
  using Set4 = Set4T<double>;
  using TSet4 = Set4T<T>;
  static constexpr int kMaxSub = 16;

  explicit Adjoint(Ctx<T>& f) : F(f) {
    check_scope();
    const auto& hp = F.hp_;
    moist_ = hp.moisture != 0;
    ns_ = std::max(1, hp.acoustic_ns);
    nsub_[0] = 1;
    nsub_[1] = std::max(1, ns_ / 2);
    nsub_[2] = ns_;
    const size_t n = F.npts_;
    auto set5 = [&](auto& s) {
      using U = std::remove_pointer_t<decltype(s.u)>;
      s.u = A.alloc<U>(n);
      s.v = A.alloc<U>(n);
      s.w = A.alloc<U>(n);
      s.th = A.alloc<U>(n);
      s.pip = A.alloc<double>(n);
      if (moist_) s.q = A.alloc<U>(n);
    };
    auto set4 = [&](auto& s) {
      using U = std::remove_pointer_t<decltype(s.a)>;
      s.a = A.alloc<U>(n);
      s.b = A.alloc<U>(n);
      s.c = A.alloc<U>(n);
      s.d = A.alloc<U>(n);
    };
    for (int m = 0; m < 3; ++m) {
      set5(est_[m]);
      set4(mf_[m]);
      set4(T_[m]);
      set5(dest_[m]);
      set4(dmf_[m]);
      set4(dT_[m]);
      for (int s = 0; s < nsub_[m]; ++s) {
        set5(sub_[m][s]);
        set5(dsub_[m][s]);
      }
    }
    set5(x0_);
    set5(xb_);
    set5(tmp_);
    zero_q_ = A.alloc<T>(n);
    pd_ = moist_ && hp.pd_moist;
    for (int m = 0; m < 3; ++m) {
      tpi_tr_[m] = A.alloc<T>(n);
      dtpi_[m] = A.alloc<double>(n);
      if (pd_) {
        ratio_tr_[m] = A.alloc<T>(n);
        dratio_[m] = A.alloc<double>(n);
      }
    }
    if (moist_) {
      qc_scr_ = A.alloc<T>(n);
      qr_scr_ = A.alloc<T>(n);
    }
    jsum_ = A.alloc<double>(1);
    wgt_ = A.alloc<double>(F.n2_);
    widen_constants();
  }

  // The linearised model covers the step without microphysics or surface
  // physics: dry, or with q_v transported and acting through buoyancy and
  // theta_v. The context can change after the adjoint
  // was set up (relaxation switched on by wfe_set_bdy_weights, the state
  // overwritten), so every entry checks again and follows the context.
  void check_scope() const {
    const auto& hp = F.hp_;
    std::string why;
    if (hp.moisture && hp.microphysics) why = "microphysics is on";
    else if (hp.diff_K < 0) why = "the diffusion coefficient is negative";
    else if (F.phys_on_) why = "the surface physics is on";
    else if (F.g_.nz + 1 > 64) why = "more than 63 levels";
    else if (std::max(1, hp.acoustic_ns) > kMaxSub) why = "more than 16 acoustic substeps";
    if (why.empty() && hp.moisture) {
      // q_c and q_r are not linearised: without microphysics they stay
      // zero when they start at zero
      std::vector<T> q(F.npts_);
      for (const T* d : {F.S_[F.cur_].qc, F.S_[F.cur_].qr}) {
        WFE_HIP(hipMemcpy(q.data(), d, F.npts_ * sizeof(T), hipMemcpyDeviceToHost));
        for (T x : q)
          if (x != 0) why = "q_c or q_r is non-zero";
      }
    }
    if (!why.empty())
      throw std::runtime_error("adjoint: the step is outside the linearised model (" + why + ")");
  }

  // Forward step of F from its current state, keeping the trajectory.
  void record_step(double dt) {
    Ctx<T>& f = F;
    hipStream_t st = f.stream_;
    if (!f.tuned_) f.tune_acoustic(dt / ns_);
    const StateSet<T> sn = f.S_[f.cur_];
    f.fill_ghosts(sn);
    copy5(est_[0], as5(sn), st);
    const double sdt[3] = {dt / 3, dt / 2, dt};
    for (int m = 0; m < 3; ++m) {
      if (m > 0) {
        copy5(est_[m], sub_[m - 1][nsub_[m - 1] - 1], st);
        f.fill_ghosts(est_set(m, sn));
      }
      // q_v of the stage = q_v,n + dts T_q(estimate), into the stage's
      // output set; q_c, q_r (zero) go to scratch, or into the new state at
      // the end. Without relaxation the update is formed where T_q is (as
      // the forward step does); with it, after the relaxation term.
      const bool bdy = f.bdy_active_;
      StateSet<T> qo{};
      if (moist_) {
        const StateSet<T>& nxt = f.S_[f.cur_ ^ 1];
        qo.qv = sub_[m][nsub_[m] - 1].q;
        qo.qc = m == 2 ? nxt.qc : qc_scr_;
        qo.qr = m == 2 ? nxt.qr : qr_scr_;
      }
      f.slow_tendencies(est_set(m, sn), dt, moist_ && !bdy ? &qo : nullptr, sdt[m], true);
      if (pd_) copy(ratio_tr_[m], f.pratio_[0], st);
      if (bdy) {
        f.bdy_relax(est_set(m, sn));
        copy(tpi_tr_[m], f.tpi_, st);
        if (moist_) f.stage_moisture(qo, sdt[m]);
      }
      copy(mf_[m].a, f.mfx_, st);
      copy(mf_[m].b, f.mfy_, st);
      copy(mf_[m].c, f.mfz_, st);
      copy(mf_[m].d, f.div_, st);
      copy(T_[m].a, f.tu_, st);
      copy(T_[m].b, f.tv_, st);
      copy(T_[m].c, f.tw_, st);
      copy(T_[m].d, f.tth_, st);
      const double dtau = sdt[m] / nsub_[m];
      for (int s = 0; s < nsub_[m]; ++s) {
        const TSet src = s == 0 ? as5(sn) : sub_[m][s - 1];
        const TSet& dst = sub_[m][s];
        dev::FusedArgs<T> a{};
        a.uv.u_src = src.u;
        a.uv.v_src = src.v;
        a.uv.u_dst = dst.u;
        a.uv.v_dst = dst.v;
        a.uv.pcur = src.pip;
        a.uv.pprev = s <= 1 ? sn.pip : sub_[m][s - 2].pip;
        a.uv.pn = sn.pip;
        a.uv.thp_cur = src.th;
        a.uv.thp_n = sn.thp;
        a.uv.qv = moist_ ? sn.qv : zero_q_;  // time-n q_v
        a.uv.tu = T_[m].a;
        a.uv.tv = T_[m].b;
        a.uv.dtau = (T)dtau;
        a.w_src = src.w;
        a.w_dst = dst.w;
        a.pip_dst = dst.pip;
        a.thp_dst = dst.th;
        a.tw = T_[m].c;
        a.tth = T_[m].d;
        a.tpi = bdy ? tpi_tr_[m] : nullptr;
        a.tpw = f.bdy_wgt_;
        f.methods_[f.method_].run(a);
      }
    }
    copy5(as5(f.S_[f.cur_ ^ 1]), sub_[2][nsub_[2] - 1], st);
    f.cur_ ^= 1;
    check_launch("adjoint record");
  }

  // dy = TL(dx) around the recorded step (device buffers).
  void tl(double dt, const Set5& dx, const Set5& dy) {
    Ctx<T>& f = F;
    dt_ = dt;
    hipStream_t st = f.stream_;
    const double sdt[3] = {dt / 3, dt / 2, dt};
    // Cells the TL kernels read but never write (the rigid-lid w levels,
    // ghost rings of the substep outputs) must hold zero, not what an earlier
    // adjoint pass left in the shared buffers.
    clear_perturbations();
    copy5(dest_[0], dx, st);
    fill5(dest_[0]);
    for (int m = 0; m < 3; ++m) {
      if (m > 0) {
        copy5(dest_[m], dsub_[m - 1][nsub_[m - 1] - 1], st);
        fill5(dest_[m]);
      }
      {
        dim3 grd(cdiv(f.g_.nx + 5, 64), cdiv(f.g_.ny + 5, 4), (unsigned)f.g_.nz);
        dev::k_massflux<double><<<grd, dim3(64, 4), 0, st>>>(
            f.g_, b64_, m64_, dest_[m].u, dest_[m].v, dest_[m].w, dmf_[m].a, dmf_[m].b,
            dmf_[m].c, dmf_[m].d);
      }
      dev::k_tl_tend<<<grid3(), dim3(64, 4), 0, st>>>(f.g_, b64_, m64_, traj(m), pert(m),
                                                       tend(m, sdt[m]));
      if (pd_) {
        dev::k_tl_pd_ratio<T><<<grid3(), dim3(64, 4), 0, st>>>(
            f.g_, b64_, m64_, dt, traj(m), pert(m), dev::Traj<T>{ratio_tr_[m]}, dratio_[m]);
        dev::k_tl_pd_apply<T><<<grid3(), dim3(64, 4), 0, st>>>(
            f.g_, f.p_, b64_, m64_, traj(m), pert(m), dev::Traj<T>{ratio_tr_[m]}, dratio_[m],
            tend(m, sdt[m]));
      }
      if (f.bdy_active_)
        dev::k_tl_relax<T><<<grid3(), dim3(64, 4), 0, st>>>(f.g_, dev::Traj<T>{f.bdy_wgt_},
                                                            pert(m), tend(m, sdt[m]));
      const double dtau = sdt[m] / nsub_[m];
      for (int s = 0; s < nsub_[m]; ++s) {
        const TSet xs = s == 0 ? est_[0] : sub_[m][s - 1];
        const Set5 ds = s == 0 ? dest_[0] : dsub_[m][s - 1];
        dev::UVTraj<T> X{xs.pip, s <= 1 ? est_[0].pip : sub_[m][s - 2].pip, est_[0].pip, xs.th,
                         est_[0].th, est_[0].q, xs.u, xs.v, T_[m].a, T_[m].b};
        dev::UVTan D{ds.u, ds.v, dsub_[m][s].u, dsub_[m][s].v, ds.pip,
                     s <= 1 ? dest_[0].pip : dsub_[m][s - 2].pip, dest_[0].pip, ds.th,
                     dest_[0].th, dT_[m].a, dT_[m].b, dest_[0].q};
        dev::k_tl_uv<<<grid_uv(), dim3(64, 4), 0, st>>>(f.g_, f.p_, b64_, m64_, X, D, dtau);
        dev::WPTraj<T> W{sub_[m][s].u, sub_[m][s].v, xs.w,       xs.pip,
                         xs.th,       est_[0].q,    tpi_tr_[m], f.bdy_active_ ? f.bdy_wgt_ : nullptr};
        dev::WPTan DW{dsub_[m][s].u, dsub_[m][s].v, ds.w, ds.pip, ds.th, dT_[m].c, dT_[m].d,
                      dsub_[m][s].w, dsub_[m][s].pip, dsub_[m][s].th, dest_[0].q, dtpi_[m]};
        dev::k_tl_wpi<T, 64><<<grid_col(), 64, 0, st>>>(f.g_, f.p_, b64_, m64_, c64_, W,
                                                       T_[m].c, DW, dtau);
      }
    }
    copy5(dy, dsub_[2][nsub_[2] - 1], st);
    check_launch("adjoint tl");
  }

  // xb = AD(yb) around the recorded step (device buffers; yb is read only).
  void ad(double dt, const Set5& yb, const Set5& xb) {
    Ctx<T>& f = F;
    dt_ = dt;
    hipStream_t st = f.stream_;
    const double sdt[3] = {dt / 3, dt / 2, dt};
    clear_perturbations();
    copy5(dsub_[2][nsub_[2] - 1], yb, st);
    for (int m = 2; m >= 0; --m) {
      const double dtau = sdt[m] / nsub_[m];
      for (int s = nsub_[m] - 1; s >= 0; --s) {
        const TSet xs = s == 0 ? est_[0] : sub_[m][s - 1];
        const Set5 ds = s == 0 ? dest_[0] : dsub_[m][s - 1];
        dev::WPTraj<T> W{sub_[m][s].u, sub_[m][s].v, xs.w,       xs.pip,
                         xs.th,       est_[0].q,    tpi_tr_[m], f.bdy_active_ ? f.bdy_wgt_ : nullptr};
        dev::WPTan AW{dsub_[m][s].u, dsub_[m][s].v, ds.w, ds.pip, ds.th, dT_[m].c, dT_[m].d,
                      dsub_[m][s].w, dsub_[m][s].pip, dsub_[m][s].th, dest_[0].q, dtpi_[m]};
        dev::k_ad_wpi<T, 64><<<grid_col(), 64, 0, st>>>(f.g_, f.p_, b64_, m64_, c64_, W,
                                                       T_[m].c, AW, dtau);
        dev::UVTraj<T> X{xs.pip, s <= 1 ? est_[0].pip : sub_[m][s - 2].pip, est_[0].pip, xs.th,
                         est_[0].th, est_[0].q, xs.u, xs.v, T_[m].a, T_[m].b};
        dev::UVTan AU{ds.u, ds.v, dsub_[m][s].u, dsub_[m][s].v, ds.pip,
                      s <= 1 ? dest_[0].pip : dsub_[m][s - 2].pip, dest_[0].pip, ds.th,
                      dest_[0].th, dT_[m].a, dT_[m].b, dest_[0].q};
        dev::k_ad_uv<<<grid_uv(), dim3(64, 4), 0, st>>>(f.g_, f.p_, b64_, m64_, X, AU, dtau);
      }
      dev::k_ad_tend<<<grid3(), dim3(64, 4), 0, st>>>(f.g_, b64_, m64_, traj(m), pert(m),
                                                       tend(m, sdt[m]));
      if (f.bdy_active_)
        dev::k_ad_relax<T><<<grid3(), dim3(64, 4), 0, st>>>(f.g_, dev::Traj<T>{f.bdy_wgt_},
                                                            pert(m), tend(m, sdt[m]));
      if (pd_) {
        dev::k_ad_pd_apply<T><<<grid3(), dim3(64, 4), 0, st>>>(
            f.g_, f.p_, b64_, m64_, traj(m), pert(m), dev::Traj<T>{ratio_tr_[m]}, dratio_[m],
            tend(m, sdt[m]));
        dev::k_ad_pd_ratio<T><<<grid3(), dim3(64, 4), 0, st>>>(
            f.g_, b64_, m64_, dt, traj(m), pert(m), dev::Traj<T>{ratio_tr_[m]}, dratio_[m]);
      }
      {
        dim3 grd(cdiv(f.g_.nx + 5, 64), cdiv(f.g_.ny + 5, 4), (unsigned)f.g_.nz);
        dev::k_ad_massflux<<<grd, dim3(64, 4), 0, st>>>(f.g_, m64_, dmf_[m].a, dmf_[m].b,
                                                        dmf_[m].c, dmf_[m].d);
        dev::k_ad_massflux_uvw<<<grd, dim3(64, 4), 0, st>>>(f.g_, b64_, m64_, dest_[m].u,
                                                            dest_[m].v, dest_[m].w, dmf_[m].a,
                                                            dmf_[m].b, dmf_[m].c);
      }
      if (m > 0) {
        ad_fill5(dest_[m]);
        add5(dsub_[m - 1][nsub_[m - 1] - 1], dest_[m]);
      }
    }
    ad_fill5(dest_[0]);
    copy5(xb, dest_[0], st);
    check_launch("adjoint ad");
  }

  // Single-step TL / AD through host arrays (reference layout), around the
  // context's current state, which is left unchanged.
  // The host entries leave the context as they found it: the saved state
  // goes back into the same state set, and the set that is current (which
  // record_step flips) is restored, so ensembles see the parity unchanged.
  void tl_step_host(double dt, const double* const* dx, double* const* dy, int nf) override {
    check_scope();
    const int par = F.cur_;
    save_state(x0_);
    record_step(dt);
    upload5(tmp_, dx, nf);
    tl(dt, tmp_, xb_);
    download5(xb_, dy, nf);
    F.cur_ = par;
    restore_state(x0_);
  }
  void ad_step_host(double dt, const double* const* yb, double* const* xb, int nf) override {
    check_scope();
    const int par = F.cur_;
    save_state(x0_);
    record_step(dt);
    upload5(tmp_, yb, nf);
    ad(dt, tmp_, xb_);
    download5(xb_, xb, nf);
    F.cur_ = par;
    restore_state(x0_);
  }

  // J = (1/N) sum_{n=1..N} sum_{i,j} wgt(i,j) 0.5 (u^2 + v^2) at the lowest
  // level, and its gradient with respect to the initial state.
  double gradient(double dt, int nsteps, int ckpt, const double* wgt2d, double* const* grad,
                  int nf) override {
    if (nsteps < 1 || ckpt < 1) throw ArgError("adjoint gradient: nsteps, ckpt must be >= 1");
    check_scope();
    Ctx<T>& f = F;
    hipStream_t st = f.stream_;
    const int par = f.cur_;
    {
      const std::vector<double> w = f.template to_dev<double>(wgt2d, false);
      WFE_HIP(hipMemcpy(wgt_, w.data(), f.n2_ * sizeof(double), hipMemcpyHostToDevice));
    }
    const double scale = 1.0 / nsteps;
    DevArena B;  // checkpoints and segment states, freed on return
    save_state(x0_);
    const int nck = (nsteps + ckpt - 1) / ckpt;
    std::vector<TSet> cks(nck), seg(ckpt);
    for (std::vector<TSet>* v : {&cks, &seg})
      for (TSet& c : *v) {
        c.u = B.alloc<T>(f.npts_);
        c.v = B.alloc<T>(f.npts_);
        c.w = B.alloc<T>(f.npts_);
        c.th = B.alloc<T>(f.npts_);
        c.pip = B.alloc<double>(f.npts_);
        if (moist_) c.q = B.alloc<T>(f.npts_);
      }
    // forward sweep: checkpoints and J
    WFE_HIP(hipMemsetAsync(jsum_, 0, sizeof(double), st));
    for (int n = 0; n < nsteps; ++n) {
      if (n % ckpt == 0) save_state(cks[n / ckpt]);
      record_step(dt);
      objective(as5(f.S_[f.cur_]), scale);
    }
    double J = 0;
    WFE_HIP(hipMemcpyAsync(&J, jsum_, sizeof(double), hipMemcpyDeviceToHost, st));
    // backward sweep
    zero5(xb_);
    for (int c = nck - 1; c >= 0; --c) {
      const int n0 = c * ckpt, n1 = std::min(nsteps, n0 + ckpt);
      restore_state(cks[c]);
      for (int n = n0; n < n1; ++n) {
        save_state(seg[n - n0]);
        record_step(dt);
      }
      for (int n = n1 - 1; n >= n0; --n) {
        // dJ/dx_{n+1} from the state after step n
        restore_state(seg[n - n0]);
        record_step(dt);
        objective_grad(as5(f.S_[f.cur_]), xb_, scale);
        copy5(tmp_, xb_, st);
        ad(dt, tmp_, xb_);
      }
    }
    download5(xb_, grad, nf);
    f.cur_ = par;
    restore_state(x0_);
    WFE_HIP(hipStreamSynchronize(st));
    return J;
  }

 private:
  // Base state, metric and acoustic coefficients for the TL/AD kernels in
  // FP64: the context's own in an FP64 context; in a mixed one, its FP32
  // values widened (the values its forward step used).
  void widen_constants() {
    if constexpr (std::is_same_v<T, double>) {
      b64_ = F.base_;
      m64_ = F.met_;
      c64_ = F.coef_;
    } else {
      const size_t n3 = F.npts_, n2 = F.n2_, n1 = (size_t)F.g_.NZ;
      auto wide = [&](const T* d, size_t n) -> const double* {
        std::vector<T> h(n);
        WFE_HIP(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
        const std::vector<double> w(h.begin(), h.end());
        double* o = A.alloc<double>(n);
        WFE_HIP(hipMemcpy(o, w.data(), n * sizeof(double), hipMemcpyHostToDevice));
        return o;
      };
      const auto& b = F.base_;
      b64_.thb = wide(b.thb, n3);
      b64_.rhob = wide(b.rhob, n3);
      b64_.dthbdz = wide(b.dthbdz, n3);
      b64_.thbw = wide(b.thbw, n3);
      b64_.rhobw = wide(b.rhobw, n3);
      b64_.qvb = wide(b.qvb, n3);
      b64_.ub = wide(b.ub, n1);
      b64_.vb = wide(b.vb, n1);
      b64_.pib = b.pib;
      b64_.hom = b.hom;
      const auto& m = F.met_;
      m64_.zeta_w = wide(m.zeta_w, n1);
      m64_.zeta_c = wide(m.zeta_c, n1);
      m64_.dzeta_c = wide(m.dzeta_c, n1);
      m64_.dzeta_w = wide(m.dzeta_w, n1);
      m64_.hx = wide(m.hx, n2);
      m64_.hy = wide(m.hy, n2);
      m64_.hxu = wide(m.hxu, n2);
      m64_.hyv = wide(m.hyv, n2);
      m64_.jac = wide(m.jac, n2);
      m64_.fcor = wide(m.fcor, n2);
      m64_.mapf = wide(m.mapf, n2);
      m64_.zt = m.zt;
      m64_.rdzc = m.rdzc;
      m64_.rdzw = m.rdzw;
      m64_.rzt = m.rzt;
      m64_.ray_c = wide(m.ray_c, n1);
      m64_.ray_w = wide(m.ray_w, n1);
      const auto& cf = F.coef_;
      c64_.rtjx = wide(cf.rtjx, n3);
      c64_.rtjy = wide(cf.rtjy, n3);
      c64_.kt3 = wide(cf.kt3, n3);
      c64_.aw = wide(cf.aw, n3);
    }
  }
  TSet as5(const StateSet<T>& s) const {
    return TSet{s.u, s.v, s.w, s.thp, s.pip, moist_ ? s.qv : nullptr};
  }
  StateSet<T> est_set(int m, const StateSet<T>& sn) const {
    StateSet<T> e = sn;
    e.u = est_[m].u;
    e.v = est_[m].v;
    e.w = est_[m].w;
    e.thp = est_[m].th;
    e.pip = est_[m].pip;
    if (moist_) e.qv = est_[m].q;
    return e;
  }
  dev::TanTraj<T> traj(int m) const {
    return dev::TanTraj<T>{est_[m].u, est_[m].v, est_[m].w, est_[m].th, mf_[m].a,
                           mf_[m].b,  mf_[m].c,  mf_[m].d,  est_[m].q};
  }
  dev::TanPert pert(int m) const {
    return dev::TanPert{dest_[m].u, dest_[m].v, dest_[m].w, dest_[m].th, dmf_[m].a,
                        dmf_[m].b,  dmf_[m].c,  dmf_[m].d,  dest_[m].q,  dest_[m].pip};
  }
  // Tendencies of stage m; moist: time-n q_v and the stage's q_v output
  // (held in the stage's last substep set).
  dev::TanTend tend(int m, double dts) const {
    const Set4& t = dT_[m];
    return dev::TanTend{t.a, t.b, t.c, t.d, dest_[0].q, dsub_[m][nsub_[m] - 1].q, dts,
                        moist_ ? 1 : 0, F.hp_.diff_K, dtpi_[m], pd_ ? 1 : 0, dt_,
                        F.hp_.w_damping ? 1 : 0};
  }
  dim3 grid3() const { return dim3(cdiv(F.g_.nx, 64), cdiv(F.g_.ny, 4), (unsigned)F.g_.nz); }
  dim3 grid_uv() const {
    return dim3(cdiv(F.g_.nx + 1, 64), cdiv(F.g_.ny + 1, 4), (unsigned)F.g_.nz);
  }
  dim3 grid_col() const { return dim3(cdiv(F.g_.nx, 64), (unsigned)F.g_.ny); }

  template <class U>
  void copy(U* d, const U* s, hipStream_t st) {
    WFE_HIP(hipMemcpyAsync(d, s, F.npts_ * sizeof(U), hipMemcpyDeviceToDevice, st));
  }
  template <class U>
  void copy5(const Set5T<U>& d, const Set5T<U>& s, hipStream_t st) {
    copy(d.u, s.u, st);
    copy(d.v, s.v, st);
    copy(d.w, s.w, st);
    copy(d.th, s.th, st);
    copy(d.pip, s.pip, st);
    if (moist_) copy(d.q, s.q, st);
  }
  void zero(double* d) {
    WFE_HIP(hipMemsetAsync(d, 0, F.npts_ * sizeof(double), F.stream_));
  }
  void clear_perturbations() {
    for (int m = 0; m < 3; ++m) {
      zero(dtpi_[m]);
      if (pd_) zero(dratio_[m]);
      zero5(dest_[m]);
      zero4(dmf_[m]);
      zero4(dT_[m]);
      for (int s = 0; s < nsub_[m]; ++s) zero5(dsub_[m][s]);
    }
  }
  void zero5(const Set5& s) {
    zero(s.u);
    zero(s.v);
    zero(s.w);
    zero(s.th);
    zero(s.pip);
    if (moist_) zero(s.q);
  }
  void zero4(const Set4& s) {
    zero(s.a);
    zero(s.b);
    zero(s.c);
    zero(s.d);
  }
  dev::FillList<double> fill_list(const Set5& s) const {
    dev::FillList<double> L{};
    const int nx = F.g_.nx, ny = F.g_.ny;
    double* f[6] = {s.u, s.v, s.w, s.th, s.pip, s.q};
    const int kind[6] = {dev::kFillScalar, dev::kFillScalar, dev::kFillW, dev::kFillScalar,
                         dev::kFillScalar, dev::kFillScalar};
    const int xl[6] = {nx, nx - 1, nx - 1, nx - 1, nx - 1, nx - 1};
    const int yl[6] = {ny - 1, ny, ny - 1, ny - 1, ny - 1, ny - 1};
    L.n = moist_ ? 6 : 5;
    for (int q = 0; q < L.n; ++q) {
      L.f[q] = f[q];
      L.kind[q] = kind[q];
      L.xlast[q] = xl[q];
      L.ylast[q] = yl[q];
    }
    return L;
  }
  void fill5(const Set5& s) {
    const dev::FillList<double> L = fill_list(s);
    const dev::FillList<double> none{};
    const int ncell = dev::ghost_count(F.g_);
    dev::k_fill_all<double><<<cdiv(ncell, 256), 256, 0, F.stream_>>>(F.g_, F.p_.bcx, F.p_.bcy, L,
                                                                     none);
  }
  void ad_fill5(const Set5& s) {
    const dev::FillList<double> L = fill_list(s);
    dev::AdFill A{};
    A.n = L.n;
    for (int q = 0; q < L.n; ++q) {
      A.f[q] = L.f[q];
      A.kind[q] = L.kind[q];
      A.xlast[q] = L.xlast[q];
      A.ylast[q] = L.ylast[q];
    }
    const int ncell = dev::ghost_count(F.g_);
    dev::k_ad_fill<<<cdiv(ncell, 256), 256, 0, F.stream_>>>(F.g_, F.p_.bcx, F.p_.bcy, A);
  }
  void add5(const Set5& d, const Set5& s) {
    const int n = (int)F.npts_;
    double* dd[6] = {d.u, d.v, d.w, d.th, d.pip, d.q};
    const double* ss[6] = {s.u, s.v, s.w, s.th, s.pip, s.q};
    for (int q = 0; q < (moist_ ? 6 : 5); ++q)
      dev::k_axpy<<<cdiv(n, 256), 256, 0, F.stream_>>>(n, 1.0, ss[q], dd[q]);
  }
  void objective(const TSet& x, double scale) {
    dev::k_ke_objective<T><<<dim3(cdiv(F.g_.nx, 64), cdiv(F.g_.ny, 4)), dim3(64, 4), 0,
                          F.stream_>>>(F.g_, x.u, x.v, wgt_, scale, jsum_);
  }
  void objective_grad(const TSet& x, const Set5& xb, double scale) {
    dev::k_ke_gradient<T><<<dim3(cdiv(F.g_.nx, 64), cdiv(F.g_.ny, 4)), dim3(64, 4), 0,
                         F.stream_>>>(F.g_, x.u, x.v, wgt_, scale, xb.u, xb.v);
  }
  void save_state(const TSet& d) { copy5(d, as5(F.S_[F.cur_]), F.stream_); }
  void restore_state(const TSet& s) { copy5(as5(F.S_[F.cur_]), s, F.stream_); }
  // nf host arrays in; q_v (the sixth) is used when moist, else ignored; a
  // moist set given five arrays gets a zero q_v.
  void upload5(const Set5& d, const double* const* h, int nf) {
    double* dd[6] = {d.u, d.v, d.w, d.th, d.pip, d.q};
    WFE_HIP(hipStreamSynchronize(F.stream_));
    for (int q = 0; q < (moist_ ? 6 : 5); ++q) {
      if (q >= nf) {
        WFE_HIP(hipMemset(dd[q], 0, F.npts_ * sizeof(double)));
        continue;
      }
      const std::vector<double> t = F.template to_dev<double>(h[q], true);
      WFE_HIP(hipMemcpy(dd[q], t.data(), F.npts_ * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  // nf host arrays out; a sixth without moisture is all zero.
  void download5(const Set5& s, double* const* h, int nf) {
    const double* ss[6] = {s.u, s.v, s.w, s.th, s.pip, s.q};
    WFE_HIP(hipStreamSynchronize(F.stream_));
    std::vector<double> t(F.npts_);
    for (int q = 0; q < nf; ++q) {
      if (q == 5 && !moist_) {
        std::fill(h[q], h[q] + F.host_n3(), 0.0);
        continue;
      }
      WFE_HIP(hipMemcpy(t.data(), ss[q], F.npts_ * sizeof(double), hipMemcpyDeviceToHost));
      F.to_host(t, h[q], true);
    }
  }

  Ctx<T>& F;
  DevArena A;
  dev::Base<double> b64_{};
  dev::Metric<double> m64_{};
  dev::Coef<double> c64_{};
  int ns_ = 1, nsub_[3] = {1, 1, 1};
  TSet est_[3], sub_[3][kMaxSub], x0_;
  Set5 dest_[3], dsub_[3][kMaxSub];
  TSet4 mf_[3], T_[3];
  Set4 dmf_[3], dT_[3];
  Set5 xb_, tmp_;
  bool moist_ = false;
  T* zero_q_ = nullptr;
  T* tpi_tr_[3] = {};         // relaxation pi' tendency of each stage
  double* dtpi_[3] = {};      // its perturbation / adjoint
  bool pd_ = false;           // q_v transported by the PD scheme
  double dt_ = 0.0;           // large step of the current TL / AD pass
  T* ratio_tr_[3] = {};       // q_v limiter ratio of each stage
  double* dratio_[3] = {};    // its perturbation / adjoint
  T *qc_scr_ = nullptr, *qr_scr_ = nullptr;
  double* jsum_ = nullptr;
  double* wgt_ = nullptr;
};

}  // namespace wfe

// ------------------------------------------------------------------ C ABI

struct wfe_ctx {
  std::unique_ptr<wfe::CtxBase> impl;
  int precision = 0;
  std::unique_ptr<wfe::AdjointBase> adj;  // created on first use
};

namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    return WFE_OK;
  } catch (const wfe::ArgError& e) {
    wfe::set_last_error(e.what());
    return WFE_E_ARG;
  } catch (const wfe::HipError& e) {
    wfe::set_last_error(e.what());
    return WFE_E_HIP;
  } catch (const std::bad_alloc&) {
    wfe::set_last_error("host allocation failed");
    return WFE_E_NOMEM;
  } catch (const std::exception& e) {
    wfe::set_last_error(e.what());
    return WFE_E_CONFIG;
  }
}

int need_ctx(wfe_ctx* c) {
  if (!c || !c->impl) {
    wfe::set_last_error("null context");
    return WFE_E_ARG;
  }
  return WFE_OK;
}

void check_n3(wfe_ctx* c, size_t n) {
  if (n != c->impl->host_n3()) throw wfe::ArgError("3-D array element count mismatch");
}
void check_n2(wfe_ctx* c, size_t n) {
  if (n != c->impl->host_n2()) throw wfe::ArgError("2-D array element count mismatch");
}

wfe::AdjointBase& adjoint_of(wfe_ctx* ctx) {
  if (!ctx->adj) {
    if (ctx->precision == WFE_PREC_FP64)
      ctx->adj = std::make_unique<wfe::Adjoint<double>>(
          *static_cast<wfe::Ctx<double>*>(ctx->impl.get()));
    else
      ctx->adj = std::make_unique<wfe::Adjoint<float>>(
          *static_cast<wfe::Ctx<float>*>(ctx->impl.get()));
  }
  return *ctx->adj;
}

}  // namespace

namespace {

int adj_tl(wfe_ctx* ctx, double dt, const double* const* dx, double* const* dy, size_t n,
           int nf, const char* who) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] {
    if (!dx || !dy) throw wfe::ArgError(std::string(who) + ": null field arrays");
    for (int q = 0; q < nf; ++q)
      if (!dx[q] || !dy[q]) throw wfe::ArgError(std::string(who) + ": null field array");
    wfe::AdjointBase& a = adjoint_of(ctx);
    check_n3(ctx, n);
    a.tl_step_host(dt, dx, dy, nf);
  });
}

int adj_ad(wfe_ctx* ctx, double dt, const double* const* yb, double* const* xb, size_t n,
           int nf, const char* who) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] {
    if (!yb || !xb) throw wfe::ArgError(std::string(who) + ": null field arrays");
    for (int q = 0; q < nf; ++q)
      if (!yb[q] || !xb[q]) throw wfe::ArgError(std::string(who) + ": null field array");
    wfe::AdjointBase& a = adjoint_of(ctx);
    check_n3(ctx, n);
    a.ad_step_host(dt, yb, xb, nf);
  });
}

int adj_gradient(wfe_ctx* ctx, double dt, int nsteps, int ckpt_every, const double* wgt2d,
                 size_t n2, double* J, double* const* grad, size_t n, int nf, const char* who) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] {
    if (!wgt2d || !J || !grad) throw wfe::ArgError(std::string(who) + ": null argument");
    for (int q = 0; q < nf; ++q)
      if (!grad[q]) throw wfe::ArgError(std::string(who) + ": null gradient array");
    wfe::AdjointBase& a = adjoint_of(ctx);
    check_n3(ctx, n);
    check_n2(ctx, n2);
    *J = a.gradient(dt, nsteps, ckpt_every, wgt2d, grad, nf);
  });
}

}  // namespace

extern "C" {

int wfe_create(const wfe_grid* g, const wfe_params* p, const wfe_base* b, const wfe_metric* m,
               int precision, wfe_ctx** out) {
  return guarded([&] {
    if (!g || !p || !b || !m || !out) throw wfe::ArgError("null argument");
    *out = nullptr;
    if (g->ng != 3) throw wfe::ArgError("ng must be 3");
    if (g->nx < 4 || g->ny < 3 || g->nz < 4)
      throw wfe::ArgError("grid too small (nx >= 4, ny >= 3, nz >= 4)");
    if (g->nz + 1 > 320)
      throw wfe::ArgError("nz too large for the column solvers (nz <= 319)");
    // Kernels address every 3-D array with a 32-bit unsigned byte offset
    // (dev::Arr), so the largest array -- an FP64 field in the device layout,
    // rows padded to a pitch of whole 32-element groups after 32 leading
    // columns -- must stay below 4 GiB: at most 2^29 padded points.
    constexpr int kA = wfe::dev::kRowAlign;
    const double pitch = (double)((kA + g->nx + g->ng + kA - 1) / kA * kA);
    const double np = pitch * (g->ny + 6) * (g->nz + 7);
    if (np * sizeof(double) >= 4294967296.0)
      throw wfe::ArgError("grid too large: the padded device grid (row pitch " +
                          std::to_string((long)pitch) + ") holds " + std::to_string((long)np) +
                          " points; an FP64 field must stay below 4 GiB (2^29 points)");
    const double* req[] = {b->thb, b->pib, b->rhob, b->dthbdz, b->thbw, b->rhobw,
                           b->thvb, b->thvbw, b->qvb, b->ub, m->zeta_w, m->zeta_c,
                           m->dzeta_c, m->dzeta_w, m->h, m->hx, m->hy, m->hx_u,
                           m->hy_v, m->jac, m->fcor, m->mapf};
    for (const double* r : req)
      if (!r) throw wfe::ArgError("missing base-state or metric array");
    if (precision != WFE_PREC_FP64 && precision != WFE_PREC_MIXED)
      throw wfe::ArgError("unknown precision");
    auto c = std::make_unique<wfe_ctx>();
    c->precision = precision;
    if (precision == WFE_PREC_FP64)
      c->impl = std::make_unique<wfe::Ctx<double>>(*g, *p, *b, *m);
    else
      c->impl = std::make_unique<wfe::Ctx<float>>(*g, *p, *b, *m);
    *out = c.release();
  });
}

int wfe_destroy(wfe_ctx* ctx) {
  delete ctx;
  return WFE_OK;
}

int wfe_set_physics(wfe_ctx* ctx, const wfe_sfc* sfc, size_t n2) {
  if (int e = need_ctx(ctx)) return e;
  if (!sfc) {
    wfe::set_last_error("wfe_set_physics: null argument");
    return WFE_E_ARG;
  }
  return guarded([&] {
    check_n2(ctx, n2);
    ctx->impl->set_physics(*sfc);
  });
}

int wfe_step(wfe_ctx* ctx, double dt, double t) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->step(dt, t); });
}

// ------------------------------------------------------------ adjoint
int wfe_tl_step(wfe_ctx* ctx, double dt, const double* const* dx, double* const* dy, size_t n) {
  return adj_tl(ctx, dt, dx, dy, n, 5, "wfe_tl_step");
}

int wfe_ad_step(wfe_ctx* ctx, double dt, const double* const* yb, double* const* xb, size_t n) {
  return adj_ad(ctx, dt, yb, xb, n, 5, "wfe_ad_step");
}

int wfe_adj_gradient(wfe_ctx* ctx, double dt, int nsteps, int ckpt_every, const double* wgt2d,
                     size_t n2, double* J, double* const* grad, size_t n) {
  return adj_gradient(ctx, dt, nsteps, ckpt_every, wgt2d, n2, J, grad, n, 5, "wfe_adj_gradient");
}

int wfe_tl_step_q(wfe_ctx* ctx, double dt, const double* const* dx, double* const* dy,
                  size_t n) {
  return adj_tl(ctx, dt, dx, dy, n, 6, "wfe_tl_step_q");
}

int wfe_ad_step_q(wfe_ctx* ctx, double dt, const double* const* yb, double* const* xb,
                  size_t n) {
  return adj_ad(ctx, dt, yb, xb, n, 6, "wfe_ad_step_q");
}

int wfe_adj_gradient_q(wfe_ctx* ctx, double dt, int nsteps, int ckpt_every,
                       const double* wgt2d, size_t n2, double* J, double* const* grad,
                       size_t n) {
  return adj_gradient(ctx, dt, nsteps, ckpt_every, wgt2d, n2, J, grad, n, 6,
                      "wfe_adj_gradient_q");
}

// ------------------------------------------------------------ ensembles
// Members advance together: one hipGraph per step (captured fork/join over
// the members' streams), so the whole ensemble is one launch per step and
// members' kernels overlap on the device. One graph per parity of the
// members' state buffers, re-captured when a time step or a member's
// generation changes.
//
// Batched substeps (switched off): members that share a storage type, the
// substep counts, moisture and boundary relaxation (and have at most 64
// levels) would form a batch: each of their acoustic substeps one
// k_acou_uv_batch and one k_acou_wpi_batch launch over all of them, with
// each member's slow phase of a stage a branch of its own between them.
// The host path walks cleanly with the HIP calls stood in on the CPU, but
// its first run on the GPU ended in a host segmentation fault inside the
// ensemble step whose cause is not found, so WFE_ENSEMBLE_BATCH=1, which
// asks for it, is refused and every member's step is one branch.
struct wfe_ensemble {
  std::vector<wfe_ctx*> m;
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<double> gdt[2];
  std::vector<unsigned> ggen[2];
  bool batching = false;
  // Per parity: the device table of the batched members' substep arguments
  // ([stage][substep][member] per batch, batches one after another).
  void* tab[2] = {nullptr, nullptr};
  size_t tab_bytes[2] = {0, 0};
  int batched_members = 0;  // in the last captured step
  ~wfe_ensemble() {
    for (hipGraphExec_t& e : exec)
      if (e) (void)hipGraphExecDestroy(e);
    for (void* t : tab)
      if (t) (void)hipFree(t);
    for (hipEvent_t e : join) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (done) (void)hipEventDestroy(done);
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

extern "C++" {

// A batch: members (indices into the ensemble) that are contexts of one
// storage type T with equal substep structure, and their launch layout.
template <class T>
struct Batch {
  std::vector<int> idx;
  std::vector<wfe::Ctx<T>*> cs;
  int ns = 0, bx = 64, kpt = 3;
  size_t lds = 0;
  unsigned uv_blocks = 0, wp_blocks = 0;
  size_t tab_off = 0;  // entries before this batch's in the table
};

// Group the members of storage type T into batches of two or more.
template <class T>
std::vector<Batch<T>> find_batches(const wfe_ensemble& E) {
  std::vector<Batch<T>> out;
  std::vector<bool> used(E.m.size(), false);
  for (size_t a = 0; a < E.m.size(); ++a) {
    auto* ca = dynamic_cast<wfe::Ctx<T>*>(E.m[a]->impl.get());
    if (used[a] || !ca || ca->levels() > 64) continue;
    Batch<T> b;
    for (size_t q = a; q < E.m.size(); ++q) {
      auto* c = dynamic_cast<wfe::Ctx<T>*>(E.m[q]->impl.get());
      if (used[q] || !c || c->levels() > 64) continue;
      if (c->stage_nsub(1) != ca->stage_nsub(1) || c->stage_nsub(2) != ca->stage_nsub(2) ||
          c->moist() != ca->moist() || c->relaxing() != ca->relaxing())
        continue;
      b.idx.push_back((int)q);
      b.cs.push_back(c);
    }
    if (b.cs.size() < 2) continue;
    for (int q : b.idx) used[q] = true;
    out.push_back(std::move(b));
  }
  return out;
}

template <class T, int BX, int KPT>
void set_batch_lds(size_t lds) {
  const void* fns[4] = {
      reinterpret_cast<const void*>(&wfe::dev::k_acou_wpi_batch<T, 64, 16, KPT, false, false>),
      reinterpret_cast<const void*>(&wfe::dev::k_acou_wpi_batch<T, 64, 16, KPT, false, true>),
      reinterpret_cast<const void*>(&wfe::dev::k_acou_wpi_batch<T, 64, 16, KPT, true, false>),
      reinterpret_cast<const void*>(&wfe::dev::k_acou_wpi_batch<T, 64, 16, KPT, true, true>)};
  for (const void* f : fns)
    WFE_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

// Lay out the batches (block ranges, LDS) and write their host table
// entries for the members' current parity.
template <class T>
void layout_batches(std::vector<Batch<T>>& bs, const double* dt,
                    std::vector<wfe::dev::BatchMember<T>>& h) {
  using wfe::cdiv;
  for (Batch<T>& b : bs) {
    wfe::Ctx<T>& lead = *b.cs[0];
    b.ns = lead.stage_nsub(2);
    b.bx = lead.uv_tile_width();
    int nzmax = 0;
    for (auto* c : b.cs) nzmax = std::max(nzmax, c->levels());
    b.kpt = nzmax <= 48 ? 3 : 4;
    b.lds = wfe::dev::WpiTile<64, 16, 4>::lds_bytes(nzmax);
    const size_t n = b.cs.size();
    b.tab_off = h.size();
    h.resize(h.size() + (size_t)3 * b.ns * n);
    std::vector<int> uv0(n), wp0(n), gx(n), gz(n);
    b.uv_blocks = b.wp_blocks = 0;
    for (size_t q = 0; q < n; ++q) {
      const wfe::dev::Grid& g = b.cs[q]->grid();
      gx[q] = cdiv(g.nx + 1, b.bx);
      gz[q] = g.nz;
      uv0[q] = (int)b.uv_blocks;
      b.uv_blocks += (unsigned)(gx[q] * cdiv(g.ny + 1, 256 / b.bx) * g.nz);
      wp0[q] = (int)b.wp_blocks;
      b.wp_blocks += (unsigned)cdiv(g.nx * g.ny, 64);
    }
    for (int st = 0; st < 3; ++st)
      for (int sub = 0; sub < lead.stage_nsub(st); ++sub)
        for (size_t q = 0; q < n; ++q) {
          wfe::dev::BatchMember<T>& M = h[b.tab_off + ((size_t)st * b.ns + sub) * n + q];
          b.cs[q]->batch_member(M, dt[b.idx[q]], st, sub);
          M.uv_begin = uv0[q];
          M.uv_gx = gx[q];
          M.uv_gz = gz[q];
          M.wp_begin = wp0[q];
        }
  }
}

// The two batched launches of one substep of batch b on stream st.
template <class T, int BX, int KPT>
void launch_batch_substep(const Batch<T>& b, const wfe::dev::BatchMember<T>* tab, hipStream_t st) {
  namespace dev = wfe::dev;
  const int n = (int)b.cs.size();
  dev::k_acou_uv_batch<T, BX><<<b.uv_blocks, dim3(BX, 256 / BX), 0, st>>>(tab, n);
  const bool moist = b.cs[0]->moist(), tpi = b.cs[0]->relaxing();
  const dim3 bt(64, 16);
  if (moist && tpi)
    dev::k_acou_wpi_batch<T, 64, 16, KPT, true, true><<<b.wp_blocks, bt, b.lds, st>>>(tab, n);
  else if (moist)
    dev::k_acou_wpi_batch<T, 64, 16, KPT, true, false><<<b.wp_blocks, bt, b.lds, st>>>(tab, n);
  else if (tpi)
    dev::k_acou_wpi_batch<T, 64, 16, KPT, false, true><<<b.wp_blocks, bt, b.lds, st>>>(tab, n);
  else
    dev::k_acou_wpi_batch<T, 64, 16, KPT, false, false><<<b.wp_blocks, bt, b.lds, st>>>(tab, n);
  wfe::check_launch("ensemble batched substep");
}

template <class T>
void batch_substep(const Batch<T>& b, const wfe::dev::BatchMember<T>* tab, hipStream_t st) {
  if (b.bx == 32)
    b.kpt == 3 ? launch_batch_substep<T, 32, 3>(b, tab, st) : launch_batch_substep<T, 32, 4>(b, tab, st);
  else
    b.kpt == 3 ? launch_batch_substep<T, 64, 3>(b, tab, st) : launch_batch_substep<T, 64, 4>(b, tab, st);
}

// Capture batch b's step: the members' slow phases as branches off the lead
// member's stream, the batched substeps on it; every member's stream ends
// joined into the ensemble stream. Parity stays unchanged.
template <class T>
void enqueue_batch(wfe_ensemble& E, const Batch<T>& b, const wfe::dev::BatchMember<T>* tab,
                   const double* dt) {
  const size_t n = b.cs.size();
  hipStream_t ls = b.cs[0]->stream();
  hipEvent_t lead = E.join[b.idx[0]];
  for (int st = 0; st < 3; ++st) {
    for (size_t q = 0; q < n; ++q) {
      wfe::Ctx<T>& c = *b.cs[q];
      const double d = dt[b.idx[q]];
      if (q > 0 && st > 0) WFE_HIP(hipStreamWaitEvent(c.stream(), lead, 0));
      if (st > 0) c.stage_post(d, st - 1);
      c.stage_pre(d, st);
    }
    for (size_t q = 1; q < n; ++q) {
      WFE_HIP(hipEventRecord(E.join[b.idx[q]], b.cs[q]->stream()));
      WFE_HIP(hipStreamWaitEvent(ls, E.join[b.idx[q]], 0));
    }
    for (int sub = 0; sub < b.cs[0]->stage_nsub(st); ++sub)
      batch_substep(b, tab + ((size_t)st * b.ns + sub) * n, ls);
    WFE_HIP(hipEventRecord(lead, ls));
  }
  for (size_t q = 0; q < n; ++q) {
    wfe::Ctx<T>& c = *b.cs[q];
    const double d = dt[b.idx[q]];
    if (q > 0) WFE_HIP(hipStreamWaitEvent(c.stream(), lead, 0));
    c.stage_post(d, 2);
    c.flip_parity();
    c.step_post(d);
    c.flip_parity();
  }
  for (size_t q = 0; q < n; ++q) {
    WFE_HIP(hipEventRecord(E.join[b.idx[q]], b.cs[q]->stream()));
    WFE_HIP(hipStreamWaitEvent(E.s, E.join[b.idx[q]], 0));
  }
}

}  // extern "C++"

void ensemble_step(wfe_ensemble& E, const double* dt, const double* t) {
  const size_t n = E.m.size();
  // The surface physics takes the hour of day from the time of each step:
  // written on each member's stream before the ensemble stream is ordered
  // after it, so the ensemble graph reads this step's value.
  for (size_t q = 0; q < n; ++q) {
    if (E.m[q]->impl->surface_physics() && !t)
      throw wfe::ArgError("ensemble member with surface physics stepped without times "
                          "(wfe_ensemble_step_at)");
    if (t) E.m[q]->impl->set_clock(dt[q], t[q]);
  }
  // Order the ensemble stream after each member's own pending work.
  for (size_t q = 0; q < n; ++q) {
    WFE_HIP(hipEventRecord(E.join[q], E.m[q]->impl->stream()));
    WFE_HIP(hipStreamWaitEvent(E.s, E.join[q], 0));
  }
  bool ready = true;
  for (wfe_ctx* c : E.m) ready = ready && c->impl->ready_for_capture();
  const int par = E.m[0]->impl->parity();
  for (wfe_ctx* c : E.m)
    if (c->impl->parity() != par) throw wfe::ArgError("ensemble members out of step");
  if (!ready) {
    // First step of a member tunes it (synchronous timing): run eagerly.
    for (size_t q = 0; q < n; ++q) E.m[q]->impl->step(dt[q], t ? t[q] : 0.0);
    return;
  }
  std::vector<unsigned> gen(n);
  for (size_t q = 0; q < n; ++q) gen[q] = E.m[q]->impl->generation();
  const std::vector<double> dts(dt, dt + n);
  if (!E.exec[par] || E.gdt[par] != dts || E.ggen[par] != gen) {
    if (E.exec[par]) {
      WFE_HIP(hipGraphExecDestroy(E.exec[par]));
      E.exec[par] = nullptr;
    }
    // Batches and their argument tables (written before the capture, once
    // no earlier replay can still read them).
    std::vector<Batch<float>> bf;
    std::vector<Batch<double>> bd;
    if (E.batching) {
      bf = find_batches<float>(E);
      bd = find_batches<double>(E);
    }
    std::vector<wfe::dev::BatchMember<float>> hf;
    std::vector<wfe::dev::BatchMember<double>> hd;
    layout_batches(bf, dt, hf);
    layout_batches(bd, dt, hd);
    const size_t bytes_f = hf.size() * sizeof(hf[0]);
    const size_t bytes = bytes_f + hd.size() * sizeof(hd[0]);
    std::vector<bool> in_batch(n, false);
    E.batched_members = 0;
    for (const auto& b : bf)
      for (int q : b.idx) in_batch[q] = true, ++E.batched_members;
    for (const auto& b : bd)
      for (int q : b.idx) in_batch[q] = true, ++E.batched_members;
    if (bytes > 0) {
      WFE_HIP(hipStreamSynchronize(E.s));
      if (E.tab_bytes[par] < bytes) {
        if (E.tab[par]) WFE_HIP(hipFree(E.tab[par]));
        E.tab[par] = nullptr;
        E.tab_bytes[par] = 0;
        WFE_HIP(hipMalloc(&E.tab[par], bytes));
        E.tab_bytes[par] = bytes;
      }
      char* t = static_cast<char*>(E.tab[par]);
      if (!hf.empty()) WFE_HIP(hipMemcpy(t, hf.data(), bytes_f, hipMemcpyHostToDevice));
      if (!hd.empty()) WFE_HIP(hipMemcpy(t + bytes_f, hd.data(), bytes - bytes_f, hipMemcpyHostToDevice));
      for (const auto& b : bf) b.kpt == 3 ? set_batch_lds<float, 64, 3>(b.lds) : set_batch_lds<float, 64, 4>(b.lds);
      for (const auto& b : bd) b.kpt == 3 ? set_batch_lds<double, 64, 3>(b.lds) : set_batch_lds<double, 64, 4>(b.lds);
    }
    auto* tf = reinterpret_cast<const wfe::dev::BatchMember<float>*>(E.tab[par]);
    auto* td = reinterpret_cast<const wfe::dev::BatchMember<double>*>(
        static_cast<const char*>(E.tab[par]) + bytes_f);
    hipGraph_t gr = nullptr;
    WFE_HIP(hipStreamBeginCapture(E.s, hipStreamCaptureModeThreadLocal));
    try {
      const wfe::CaptureScope cap;
      WFE_HIP(hipEventRecord(E.fork, E.s));
      for (size_t q = 0; q < n; ++q) {
        hipStream_t ms = E.m[q]->impl->stream();
        WFE_HIP(hipStreamWaitEvent(ms, E.fork, 0));
        if (in_batch[q]) continue;
        E.m[q]->impl->enqueue_step(dt[q]);
        WFE_HIP(hipEventRecord(E.join[q], ms));
        WFE_HIP(hipStreamWaitEvent(E.s, E.join[q], 0));
      }
      for (const auto& b : bf) enqueue_batch(E, b, tf + b.tab_off, dt);
      for (const auto& b : bd) enqueue_batch(E, b, td + b.tab_off, dt);
    } catch (...) {
      (void)hipStreamEndCapture(E.s, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    WFE_HIP(hipStreamEndCapture(E.s, &gr));
    const hipError_t e = hipGraphInstantiate(&E.exec[par], gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      E.exec[par] = nullptr;
      throw wfe::HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    E.gdt[par] = dts;
    E.ggen[par] = gen;
  }
  WFE_HIP(hipGraphLaunch(E.exec[par], E.s));
  // Later work on a member's own stream follows the ensemble step.
  WFE_HIP(hipEventRecord(E.done, E.s));
  for (wfe_ctx* c : E.m) {
    WFE_HIP(hipStreamWaitEvent(c->impl->stream(), E.done, 0));
    c->impl->flip_parity();
  }
}

}  // namespace

int wfe_ensemble_create(wfe_ctx* const* members, int n, wfe_ensemble** out) {
  if (!members || n < 1 || !out) {
    wfe::set_last_error("wfe_ensemble_create: need members, n >= 1 and out");
    return WFE_E_ARG;
  }
  for (int q = 0; q < n; ++q) {
    if (int e = need_ctx(members[q])) return e;
    for (int r = 0; r < q; ++r)
      if (members[r] == members[q]) {
        wfe::set_last_error("wfe_ensemble_create: a context appears twice");
        return WFE_E_ARG;
      }
  }
  return guarded([&] {
    auto E = std::make_unique<wfe_ensemble>();
    E->m.assign(members, members + n);
    WFE_HIP(hipStreamCreate(&E->s));
    WFE_HIP(hipEventCreateWithFlags(&E->fork, hipEventDisableTiming));
    WFE_HIP(hipEventCreateWithFlags(&E->done, hipEventDisableTiming));
    E->join.assign(n, nullptr);
    for (int q = 0; q < n; ++q) WFE_HIP(hipEventCreateWithFlags(&E->join[q], hipEventDisableTiming));
    if (const char* e = std::getenv("WFE_ENSEMBLE_BATCH"))
      if (std::string(e) != "0")
        throw wfe::ArgError(
            "WFE_ENSEMBLE_BATCH=1: batched ensemble substeps are switched off (their first GPU "
            "run ended in a host segmentation fault of unknown cause)");
    *out = E.release();
  });
}

int wfe_ensemble_destroy(wfe_ensemble* e) {
  if (!e) return WFE_OK;
  return guarded([&] {
    (void)hipStreamSynchronize(e->s);
    delete e;
  });
}

int wfe_ensemble_step(wfe_ensemble* e, const double* dt) {
  if (!e || !dt) {
    wfe::set_last_error("wfe_ensemble_step: null argument");
    return WFE_E_ARG;
  }
  return guarded([&] { ensemble_step(*e, dt, nullptr); });
}

int wfe_ensemble_step_at(wfe_ensemble* e, const double* dt, const double* t) {
  if (!e || !dt || !t) {
    wfe::set_last_error("wfe_ensemble_step_at: null argument");
    return WFE_E_ARG;
  }
  return guarded([&] { ensemble_step(*e, dt, t); });
}

int wfe_ensemble_batched(const wfe_ensemble* e, int* members) {
  if (!e || !members) {
    wfe::set_last_error("wfe_ensemble_batched: null argument");
    return WFE_E_ARG;
  }
  *members = e->batched_members;
  return WFE_OK;
}

int wfe_ensemble_sync(wfe_ensemble* e) {
  if (!e) {
    wfe::set_last_error("wfe_ensemble_sync: null ensemble");
    return WFE_E_ARG;
  }
  return guarded([&] {
    WFE_HIP(hipStreamSynchronize(e->s));
    for (wfe_ctx* c : e->m) WFE_HIP(hipStreamSynchronize(c->impl->stream()));
  });
}

int wfe_upload(wfe_ctx* ctx, int field, const double* host, size_t n) {
  if (int e = need_ctx(ctx)) return e;
  if (!host) return WFE_E_ARG;
  return guarded([&] { ctx->impl->upload(field, host, n); });
}

int wfe_download(wfe_ctx* ctx, int field, double* host, size_t n) {
  if (int e = need_ctx(ctx)) return e;
  if (!host) return WFE_E_ARG;
  return guarded([&] { ctx->impl->download(field, host, n); });
}

int wfe_absmax(wfe_ctx* ctx, int field, float* out) {
  if (int e = need_ctx(ctx)) return e;
  if (!out) return WFE_E_ARG;
  return guarded([&] { *out = ctx->impl->absmax(field); });
}

int wfe_diagnose(wfe_ctx* ctx, wfe_diag* out) {
  if (int e = need_ctx(ctx)) return e;
  if (!out) {
    wfe::set_last_error("wfe_diagnose: null output");
    return WFE_E_ARG;
  }
  return guarded([&] {
    double r[wfe::dev::kDgSlots];
    ctx->impl->diagnose(r);
    out->wmax = r[wfe::dev::kDgW];
    out->thmin = r[wfe::dev::kDgThMin];
    out->thmax = r[wfe::dev::kDgThMax];
    out->pipmax = r[wfe::dev::kDgPip];
    out->qcmax = r[wfe::dev::kDgQc];
    out->qrmax = r[wfe::dev::kDgQr];
    out->finite = r[wfe::dev::kDgFinite] != 0.0;
  });
}

int wfe_pack_field_async(wfe_ctx* ctx, int field, float* dst, size_t n) {
  if (int e = need_ctx(ctx)) return e;
  if (!dst) {
    wfe::set_last_error("wfe_pack_field_async: null destination");
    return WFE_E_ARG;
  }
  return guarded([&] { ctx->impl->pack_async(field, dst, n); });
}

int wfe_output_mark(wfe_ctx* ctx, long long* ticket) {
  if (int e = need_ctx(ctx)) return e;
  if (!ticket) {
    wfe::set_last_error("wfe_output_mark: null ticket");
    return WFE_E_ARG;
  }
  return guarded([&] { *ticket = ctx->impl->output_mark(); });
}

int wfe_output_wait(wfe_ctx* ctx, long long ticket) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->output_wait(ticket); });
}

int wfe_device_alloc(size_t bytes, void** out) {
  return guarded([&] {
    if (!out || bytes == 0) throw wfe::ArgError("wfe_device_alloc: bad arguments");
    *out = nullptr;
    WFE_HIP(hipMalloc(out, bytes));
  });
}

int wfe_device_free(void* p) {
  return guarded([&] {
    if (p) WFE_HIP(hipFree(p));
  });
}

int wfe_device_upload(void* dev, const void* host, size_t bytes) {
  return guarded([&] {
    if (!dev || !host) throw wfe::ArgError("wfe_device_upload: null pointer");
    WFE_HIP(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
  });
}

int wfe_device_download(void* host, const void* dev, size_t bytes) {
  return guarded([&] {
    if (!dev || !host) throw wfe::ArgError("wfe_device_download: null pointer");
    WFE_HIP(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
  });
}

int wfe_host_alloc(size_t bytes, void** out) {
  return guarded([&] {
    if (!out || bytes == 0) throw wfe::ArgError("wfe_host_alloc: bad arguments");
    *out = nullptr;
    WFE_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  });
}

int wfe_host_free(void* p) {
  return guarded([&] {
    if (p) WFE_HIP(hipHostFree(p));
  });
}

int wfe_sync(wfe_ctx* ctx) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->sync(); });
}

int wfe_set_bdy_weights(wfe_ctx* ctx, const double* wgt, size_t n) {
  if (int e = need_ctx(ctx)) return e;
  if (!wgt) return WFE_E_ARG;
  return guarded([&] { ctx->impl->set_bdy_weights(wgt, n); });
}

int wfe_set_bdy_target(wfe_ctx* ctx, int slot, int var, const double* host, size_t n) {
  if (int e = need_ctx(ctx)) return e;
  if (!host) return WFE_E_ARG;
  return guarded([&] { ctx->impl->set_bdy_target(slot, var, host, n); });
}

int wfe_swap_bdy_targets(wfe_ctx* ctx) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->swap_bdy_targets(); });
}

int wfe_set_bdy_tfrac(wfe_ctx* ctx, double tfrac) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->set_bdy_tfrac(tfrac); });
}

int wfe_apply_bcs(wfe_ctx* ctx) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->apply_bcs(); });
}

int wfe_compute_tendencies(wfe_ctx* ctx, double dt) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->compute_tendencies(dt); });
}

int wfe_acoustic_substeps(wfe_ctx* ctx, double dtau, int nsub) {
  if (int e = need_ctx(ctx)) return e;
  if (nsub < 1) return WFE_E_ARG;
  return guarded([&] { ctx->impl->acoustic_substeps(dtau, nsub); });
}

int wfe_kessler(wfe_ctx* ctx, double dt) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->kessler(dt); });
}

int wfe_kernel_timing(wfe_ctx* ctx, int mode) {
  if (int e = need_ctx(ctx)) return e;
  return guarded([&] { ctx->impl->kernel_timing(mode); });
}

int wfe_kernel_bench(wfe_ctx* ctx, const char* kernel, int reps, double* avg_us) {
  if (int e = need_ctx(ctx)) return e;
  if (!kernel || !avg_us) {
    wfe::set_last_error("wfe_kernel_bench: null argument");
    return WFE_E_ARG;
  }
  return guarded([&] { *avg_us = ctx->impl->kernel_bench(kernel, reps); });
}

int wfe_kernel_timing_report(wfe_ctx* ctx, char* buf, size_t n) {
  if (int e = need_ctx(ctx)) return e;
  if (!buf || n == 0) return WFE_E_ARG;
  return guarded([&] {
    const std::string r = ctx->impl->kernel_report();
    if (r.size() + 1 > n) throw wfe::ArgError("report buffer too small");
    std::memcpy(buf, r.c_str(), r.size() + 1);
  });
}

int wfe_copy_probe(size_t bytes, int reps, double* gbps) {
  return guarded([&] {
    if (!gbps || reps < 1 || bytes < 4096) throw wfe::ArgError("copy probe: bad arguments");
    const size_t n4 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    WFE_HIP(hipMalloc(&a, n4 * 16));
    WFE_HIP(hipMalloc(&b, n4 * 16));
    struct Free {
      void *a, *b;
      ~Free() {
        (void)hipFree(a);
        (void)hipFree(b);
      }
    } fr{a, b};
    WFE_HIP(hipMemset(a, 0, n4 * 16));
    hipStream_t st;
    WFE_HIP(hipStreamCreate(&st));
    hipEvent_t e0, e1;
    WFE_HIP(hipEventCreate(&e0));
    WFE_HIP(hipEventCreate(&e1));
    const unsigned blocks = 256 * 8;
    wfe::dev::k_copy_probe<<<blocks, 256, 0, st>>>(static_cast<const float4*>(a),
                                                    static_cast<float4*>(b), n4);
    WFE_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r)
      wfe::dev::k_copy_probe<<<blocks, 256, 0, st>>>(static_cast<const float4*>(a),
                                                      static_cast<float4*>(b), n4);
    WFE_HIP(hipEventRecord(e1, st));
    WFE_HIP(hipEventSynchronize(e1));
    float ms = 0;
    WFE_HIP(hipEventElapsedTime(&ms, e0, e1));
    wfe::check_launch("copy probe");
    *gbps = 2.0 * (double)(n4 * 16) * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
  });
}

const char* wfe_last_error(void) { return wfe::g_last_error.c_str(); }

int wfe_abi_version(void) { return WFE_HIP_ABI_VERSION; }

int wfe_device_name(char* buf, size_t n) {
  return guarded([&] {
    if (!buf || n == 0) throw wfe::ArgError("null buffer");
    int dev = 0;
    WFE_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop{};
    WFE_HIP(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, n, "%s", prop.name);
  });
}

}  // extern "C"
