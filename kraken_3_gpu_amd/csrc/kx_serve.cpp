// Serving loop: continuous batching of ES rollout requests on one GPU.
//
// Requests arrive on any thread (kx_server_submit). An admission thread takes
// them in arrival order while a slot is free and free HBM (hipMemGetInfo)
// covers the request with the reserve left over; it builds the request's world
// (2*pairs creatures of one morphology at one pose on the server's map) into a
// handle of its own, uploads the initial state and sets up its policy. A
// scheduler thread runs every active request in chunks of K steps: for each
// chunk the host evaluates each request's policy for its next steps into a
// control tape (a worker pool; this overlaps the device running the previous
// chunk), the tapes go up on a copy stream, and ONE persistent launch covers
// the workgroups of every active request, each for its own remaining steps
// (kx_k_persist_batch). Requests that arrive meanwhile join at the next chunk
// boundary; finished ones are downloaded on a third stream, reported, and
// their memory returns to admission.
//
// A request's result does not depend on what else is in the batch: every
// creature's arithmetic runs in its own wavefront from its own handle's state,
// and its controls come from its own tape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/kxpbd.h"
#include "kx_internal.h"

namespace {

// ── policy: 64 -> 32 -> 48, tanh ────────────────────────────────────────
constexpr int NI = 64, NH = 32, NO = 48;

// mulberry32 (src/world/map.js:11-20)
struct Mb32 {
  uint32_t a;
  uint32_t next_u32() {
    a += 0x6D2B79F5u;
    uint32_t t = a;
    t = (t ^ (t >> 15)) * (t | 1u);
    t ^= t + (t ^ (t >> 7)) * (t | 61u);
    return t ^ (t >> 14);
  }
  double next() { return (double)next_u32() / 4294967296.0; }
};
// Box-Muller as gauss (src/brain/net.js:56-62)
double gauss(Mb32 &r) {
  double u = 0, v = 0;
  while (u == 0) u = r.next();
  while (v == 0) v = r.next();
  return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * v);
}
uint32_t pair_seed(uint32_t noise_seed, uint32_t p) {
  Mb32 r{noise_seed ^ (p * 0x9E3779B9u)};
  r.next_u32();
  return r.next_u32();
}

// tanh in binary32, the same in every loop and thread: odd Taylor series
// below 1/16, else 1 - 2/(exp(2|x|) + 1) with exp from a Cody-Waite range
// reduction by ln 2 and its degree-6 Taylor polynomial; |x| >= 9 gives +-1.
__attribute__((always_inline)) inline float tanh32(float x) {
  const float ax = std::fabs(x);
  const float x2 = x * x;
  const float small = x * (1.0f + x2 * (-1.0f / 3.0f + x2 * (2.0f / 15.0f + x2 * (-17.0f / 315.0f))));
  const float y = std::min(2.0f * ax, 18.0f);
  const float k = std::nearbyint(y * 1.44269504088896340736f);
  const float r = (y - k * 0.693145751953125f) - k * 1.42860682030941723212e-6f;
  const float p = 1.0f + r * (1.0f + r * (0.5f + r * (1.0f / 6.0f + r * (1.0f / 24.0f + r * (1.0f / 120.0f + r * (1.0f / 720.0f))))));
  uint32_t bits = (uint32_t)((int)k + 127) << 23;
  float scale;
  std::memcpy(&scale, &bits, 4);
  const float e = p * scale;
  float big = 1.0f - 2.0f / (e + 1.0f);
  big = ax >= 9.0f ? 1.0f : big;
  big = std::copysign(big, x);
  return ax < 0.0625f ? small : big;
}

// open-loop inputs: sin and cos of 32 frequencies (0.125 .. 4 Hz) at the
// step's time
void features(double t, float *x) {
  for (int k = 0; k < NI / 2; k++) {
    const double w = 2.0 * M_PI * 0.125 * (k + 1) * t;
    x[2 * k] = (float)std::sin(w);
    x[2 * k + 1] = (float)std::cos(w);
  }
}

// Policy of one request, laid out for loops that vectorise across outputs:
// W1T[i][j] (theta, shared), b1; per pair sE1T[i][j] = sigma*eps W1 part and
// sE1b; per creature W2T[j][o] and b2 (theta +- sigma*eps, rounded to f32 as
// es.js's Float32Array genomes are).
struct Policy {
  int pairs = 0, arms = 0;
  std::vector<float> W1T, b1;          // [NI][NH], [NH]
  std::vector<float> sE1T, sE1b;       // [pairs][NI][NH], [pairs][NH]
  std::vector<float> W2T, b2;          // [2*pairs][NH][NO], [2*pairs][NO]
};

// theta in net.js's order (W1[NH][NI], b1, W2[NO][NH], b2; randomGenome's
// 1/sqrt(fan_in) scale and zero biases, src/brain/net.js:41-53)
void draw_theta(uint32_t seed, std::vector<float> &th) {
  th.assign(NI * NH + NH + NH * NO + NO, 0.0f);
  Mb32 r{seed};
  const double s1 = 1.0 / std::sqrt((double)NI), s2 = 1.0 / std::sqrt((double)NH);
  size_t i = 0;
  for (; i < (size_t)NI * NH; i++) th[i] = (float)(gauss(r) * s1);
  i += NH;
  for (size_t k = 0; k < (size_t)NH * NO; k++, i++) th[i] = (float)(gauss(r) * s2);
}
void draw_eps(uint32_t noise_seed, int p, std::vector<float> &e) {
  e.resize(NI * NH + NH + NH * NO + NO);
  Mb32 r{pair_seed(noise_seed, (uint32_t)p)};
  for (auto &v : e) v = (float)gauss(r);
}

void build_policy(const kx_request &q, Policy &P) {
  P.pairs = q.pairs;
  P.arms = q.arms;
  std::vector<float> th, e;
  draw_theta(q.policy_seed, th);
  const float *W1 = th.data(), *B1 = W1 + NI * NH, *W2 = B1 + NH, *B2 = W2 + NH * NO;
  P.W1T.resize(NI * NH);
  for (int j = 0; j < NH; j++)
    for (int i = 0; i < NI; i++) P.W1T[i * NH + j] = W1[j * NI + i];
  P.b1.assign(B1, B1 + NH);
  P.sE1T.resize((size_t)q.pairs * NI * NH);
  P.sE1b.resize((size_t)q.pairs * NH);
  P.W2T.resize((size_t)2 * q.pairs * NH * NO);
  P.b2.resize((size_t)2 * q.pairs * NO);
  for (int p = 0; p < q.pairs; p++) {
    draw_eps(q.noise_seed, p, e);
    const float *E1 = e.data(), *EB1 = E1 + NI * NH, *E2 = EB1 + NH, *EB2 = E2 + NH * NO;
    float *s1 = &P.sE1T[(size_t)p * NI * NH];
    for (int j = 0; j < NH; j++)
      for (int i = 0; i < NI; i++) s1[i * NH + j] = (float)(q.sigma * E1[j * NI + i]);
    for (int j = 0; j < NH; j++) P.sE1b[(size_t)p * NH + j] = (float)(q.sigma * EB1[j]);
    for (int sgn = 0; sgn < 2; sgn++) {
      const double s = sgn ? -q.sigma : q.sigma;
      const int c = 2 * p + sgn;
      float *w2 = &P.W2T[(size_t)c * NH * NO];
      for (int o = 0; o < NO; o++)
        for (int j = 0; j < NH; j++) w2[j * NO + o] = (float)((double)W2[o * NH + j] + (double)E2[o * NH + j] * s);
      for (int o = 0; o < NO; o++) P.b2[(size_t)c * NO + o] = (float)((double)B2[o] + (double)EB2[o] * s);
    }
  }
}

// Tape records of pairs [p0, p1) for steps step0 .. step0+n-1 into
// tape[(row * ncr + creature) * KX_TAPE_WORDS]; `A` holds W1 x + b1 per row.
__attribute__((always_inline)) inline void policy_rows_body(const Policy &P, int p0, int p1, int step0, int n,
                                                            double step_s, float *tape) {
  const int ncr = 2 * P.pairs;
  alignas(64) float x[NI], A[NH], D[NH], h[2][NH], out[NO];
  for (int r = 0; r < n; r++) {
    features((double)(step0 + r) * step_s, x);
    for (int j = 0; j < NH; j++) A[j] = P.b1[j];
    for (int i = 0; i < NI; i++) {
      const float xi = x[i];
      const float *w = &P.W1T[i * NH];
      for (int j = 0; j < NH; j++) A[j] += xi * w[j];
    }
    for (int p = p0; p < p1; p++) {
      const float *e = &P.sE1T[(size_t)p * NI * NH];
      for (int j = 0; j < NH; j++) D[j] = P.sE1b[(size_t)p * NH + j];
      for (int i = 0; i < NI; i++) {
        const float xi = x[i];
        const float *w = e + i * NH;
        for (int j = 0; j < NH; j++) D[j] += xi * w[j];
      }
      for (int j = 0; j < NH; j++) {
        h[0][j] = tanh32(A[j] + D[j]);
        h[1][j] = tanh32(A[j] - D[j]);
      }
      for (int sgn = 0; sgn < 2; sgn++) {
        const int c = 2 * p + sgn;
        const float *w2 = &P.W2T[(size_t)c * NH * NO];
        for (int o = 0; o < NO; o++) out[o] = P.b2[(size_t)c * NO + o];
        for (int j = 0; j < NH; j++) {
          const float hj = h[sgn][j];
          const float *w = w2 + j * NO;
          for (int o = 0; o < NO; o++) out[o] += hj * w[o];
        }
        for (int o = 0; o < NO; o++) out[o] = tanh32(out[o]);
        // decode (src/brain/net.js:114-124)
        float *rec = tape + ((size_t)r * ncr + c) * KX_TAPE_WORDS;
        std::memset(rec, 0, KX_TAPE_WORDS * sizeof(float));
        rec[0] = 0.5f + 0.5f * out[0];
        for (int a = 0; a < P.arms && a < KX_MAX_ARMS; a++) {
          rec[1 + 3 * a] = out[1 + 3 * a];
          rec[2 + 3 * a] = out[2 + 3 * a];
          rec[3 + 3 * a] = out[3 + 3 * a] > 0.0f ? out[3 + 3 * a] : 0.0f;
        }
      }
    }
  }
}

// The same loops compiled twice: for AVX2 (8-wide, and vector round-to-
// nearest for tanh's range reduction) and for the x86-64 baseline, chosen at
// run time. No FMA in either: every product and sum is rounded on its own, so
// both give the same bits.
__attribute__((target("avx2"))) void policy_rows_avx2(const Policy &P, int p0, int p1, int step0, int n,
                                                      double step_s, float *tape) {
  policy_rows_body(P, p0, p1, step0, n, step_s, tape);
}
void policy_rows_base(const Policy &P, int p0, int p1, int step0, int n, double step_s, float *tape) {
  policy_rows_body(P, p0, p1, step0, n, step_s, tape);
}
void policy_rows(const Policy &P, int p0, int p1, int step0, int n, double step_s, float *tape) {
  static const bool avx2 = [] {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2") != 0;
  }();
  if (avx2) policy_rows_avx2(P, p0, p1, step0, n, step_s, tape);
  else policy_rows_base(P, p0, p1, step0, n, step_s, tape);
}

// ── worker pool ────────────────────────────────────────────────────────
// Items are claimed under the lock and tagged with the run's generation, so a
// worker still leaving one run never takes an item of the next.
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 0; i < n; i++) th_.emplace_back([this] { work(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  // runs fn(0..n-1) on the workers and the caller; returns when all are done
  void run(int n, const std::function<void(int)> &fn) {
    if (n <= 0) return;
    uint64_t g;
    {
      std::lock_guard<std::mutex> l(m_);
      fn_ = &fn;
      n_ = n;
      next_ = 0;
      left_ = n;
      g = ++gen_;
    }
    cv_.notify_all();
    drain(g);
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return left_ == 0; });
    fn_ = nullptr;
  }

 private:
  void drain(uint64_t g) {
    for (;;) {
      int i;
      const std::function<void(int)> *f;
      {
        std::lock_guard<std::mutex> l(m_);
        if (gen_ != g || fn_ == nullptr || next_ >= n_) return;
        i = next_++;
        f = fn_;
      }
      (*f)(i);
      std::lock_guard<std::mutex> l(m_);
      if (--left_ == 0) done_.notify_all();
    }
  }
  void work() {
    uint64_t seen = 0;
    for (;;) {
      uint64_t g;
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return stop_ || (gen_ != seen && fn_ != nullptr); });
        if (stop_) return;
        g = seen = gen_;
      }
      drain(g);
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *fn_ = nullptr;
  int n_ = 0, next_ = 0, left_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// pinned host / device buffers of one request
struct Req {
  kx_request q{};
  double t_submit = 0, t_admit = -1, t_done = 0;
  kx_handle *h = nullptr;
  KxHandleView v{};
  Policy pol;
  int ncr = 0, nb = 0, bpc = 0;
  int launched = 0;                 // steps issued to the device
  int64_t last_chunk = -1;          // chunk that ran its last steps
  float *tape_host[2] = {nullptr, nullptr};   // pinned, [K][ncr][KX_TAPE_WORDS]
  float *tape_dev[2] = {nullptr, nullptr};
  std::vector<float> px0;           // first body x of every creature at spawn
  int status = 0;
  size_t dev_bytes = 0;
  int cost = 0;                     // relative time of one of its workgroups per step
};

struct Chunk {
  int64_t id = 0;
  hipEvent_t beg = nullptr, end = nullptr, copied = nullptr;
  std::vector<Req *> members;
  std::vector<int> nsteps;
};

double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct kx_server {
  kx_server_opts o{};
  int K = 50;
  std::chrono::steady_clock::time_point t0;
  std::unique_ptr<Pool> pool;
  hipStream_t comp = nullptr, copy = nullptr, dl = nullptr;
  int device = 0;

  std::mutex m;
  std::condition_variable cv_admit, cv_sched, cv_done;
  std::deque<Req *> queued;         // submitted, not yet admitted
  std::deque<Req *> ready;          // admitted (handle built), not yet scheduled
  std::vector<Req *> active;        // scheduled (scheduler thread only)
  std::deque<kx_request_done> done;
  int n_admitted_live = 0;          // building + ready + active
  int64_t pending_bytes = 0;        // admission estimates of requests still being built
  std::atomic<int> fold{0};         // collision fold of every handle (0 = not yet chosen, 1 parallel, 2 owner)
  bool stop = false, sched_exit = false;
  std::atomic<int64_t> launches{0}, creature_steps{0}, peak_active{0}, admit_waits{0};
  std::atomic<int> n_done{0};
  double busy = 0;

  // per-chunk staging (two chunks in flight)
  KxBatchEnt *ent_host[2] = {nullptr, nullptr};
  int2 *map_host[2] = {nullptr, nullptr};
  KxBatchEnt *ent_dev[2] = {nullptr, nullptr};
  int2 *map_dev[2] = {nullptr, nullptr};
  size_t ent_cap = 0, map_cap = 0;
  std::vector<float> dl_scratch;
  float *dl_host = nullptr;
  size_t dl_cap = 0;

  std::vector<std::thread> th_admit;  // parallel builders, taking requests in arrival order
  std::thread th_sched, th_reap;
  std::deque<Req *> reap;            // finished requests whose memory is still held
  std::condition_variable cv_reap;
  bool reap_stop = false;
  std::atomic<int> n_reaped{0};

  void admit_loop();
  void sched_loop();
  void reap_loop();
  int build(Req *r);
  void release(Req *r);
  int issue(int64_t id, Chunk &c);
  void retire(Chunk &c);
  void finish(Req *r, int status);
};

static int server_fail(int code, const char *msg) { return kx_error(code, msg); }

// ── admission ─────────────────────────────────────────────────────────
int kx_server::build(Req *r) {
  const kx_request &q = r->q;
  kx_builder *B = kx_builder_new();
  if (!B) return server_fail(KX_ENOMEM, "builder allocation failed");
  int rc = kx_builder_map(B, o.map_width, o.map_height, o.map_cols, o.map_rows, o.map_density, o.map_wall,
                          o.map_seed);
  for (int c = 0; c < 2 * q.pairs && rc >= 0; c++)
    rc = kx_builder_spawn(B, q.x, q.y, q.angle, q.arms, q.segs, q.ring, 0.28);
  kx_world_desc desc{};
  kx_creature_desc cr{};
  if (rc >= 0) rc = kx_builder_desc(B, &desc, &cr);
  if (rc < 0) {
    kx_builder_free(B);
    return rc;
  }
  const int nb = desc.n_bodies, nc = desc.n_constraints;
  r->ncr = cr.n_creatures;
  r->nb = nb;
  r->bpc = nb / std::max(r->ncr, 1);
  std::vector<double> px(nb), py(nb), ang(nb), lbx(nc), lby(nc), comp(nc), tgt(nc), bf(nc);
  std::vector<uint8_t> broken(nc);
  kx_builder_state(B, px.data(), py.data(), ang.data());
  kx_builder_cons_init(B, lbx.data(), lby.data(), comp.data(), tgt.data(), bf.data(), broken.data());
  kx_params p = o.params;
  p.path = KX_PATH_PERSISTENT;
  // lanes per creature as the bucket tuner settles them on the uneven batch
  // of the same morphologies: 8 for the smallest (at most 68 constraints,
  // whose rounds then fuse with a pin on each lane), 16 for the rest
  p.lanes_per_creature = nc <= 68 * std::max(r->ncr, 1) ? 8 : 16;
  p.waves_per_block = 1;
  rc = kx_create(&desc, &cr, &p, &r->h);
  kx_builder_free(B);
  if (rc < 0) return rc;
  // initial state (gpuworld.js upload 250-298: f64 -> f32, constraints in handle order)
  std::vector<int32_t> order(nc), color(nc);
  kx_constraint_order(r->h, order.data(), color.data());
  std::vector<float> f[6];
  for (auto &a : f) a.assign(nb, 0.0f);
  for (int i = 0; i < nb; i++) { f[0][i] = (float)px[i]; f[1][i] = (float)py[i]; f[2][i] = (float)ang[i]; }
  std::vector<float> c_lbx(nc), c_lby(nc), c_comp(nc), c_tgt(nc), c_bf(nc), zero(nc, 0.0f);
  std::vector<uint8_t> c_br(nc);
  for (int k = 0; k < nc; k++) {
    const int i = order[k];
    c_lbx[k] = (float)lbx[i]; c_lby[k] = (float)lby[i]; c_comp[k] = (float)comp[i]; c_tgt[k] = (float)tgt[i];
    c_bf[k] = std::isfinite(bf[i]) ? (float)std::min(bf[i], 1e30) : 1e30f;
    c_br[k] = broken[i] ? 1 : 0;
  }
  kx_body_state bs{};
  bs.px = f[0].data(); bs.py = f[1].data(); bs.ang = f[2].data();
  bs.vx = f[3].data(); bs.vy = f[4].data(); bs.av = f[5].data();
  kx_cons_state cs{};
  cs.lbx = c_lbx.data(); cs.lby = c_lby.data(); cs.compliance = c_comp.data(); cs.target = c_tgt.data();
  cs.break_force = c_bf.data(); cs.broken = c_br.data();
  cs.lambda = zero.data(); cs.lambda_x = zero.data(); cs.lambda_y = zero.data();
  if ((rc = kx_upload(r->h, &bs, &cs)) < 0) return rc;
  std::vector<double> cd((size_t)r->ncr * KX_MAX_ARMS, 0.0);
  if ((rc = kx_set_cooldown(r->h, cd.data())) < 0) return rc;
  if ((rc = kx_handle_view(r->h, &r->v)) < 0) return rc;
  // one batched launch runs one collision fold: the first handle built
  // chooses it (every request shares the map), later ones follow
  {
    const int mine = r->v.P.coll_par ? 1 : 2;
    int want = 0;
    if (!fold.compare_exchange_strong(want, mine)) {
      if (want != mine) {
        if ((rc = kx_set_coll_mode(r->h, want)) < 0) return rc;
        if ((rc = kx_handle_view(r->h, &r->v)) < 0) return rc;
      }
    }
  }
  if (r->v.wpb != 1) return server_fail(KX_EINVAL, "request handle is not on the 1-wave shape");
  r->px0.resize(r->ncr);
  for (int c = 0; c < r->ncr; c++) r->px0[c] = f[0][(size_t)c * r->bpc];
  // control tapes: two chunk slots, host side pinned
  const size_t tw = (size_t)K * r->ncr * KX_TAPE_WORDS;
  for (int s = 0; s < 2; s++) {
    if (hipHostMalloc((void **)&r->tape_host[s], tw * 4, hipHostMallocDefault) != hipSuccess)
      return server_fail(KX_ENOMEM, "pinned tape allocation failed");
    if (hipMalloc((void **)&r->tape_dev[s], tw * 4) != hipSuccess)
      return server_fail(KX_ENOMEM, "device tape allocation failed");
  }
  r->dev_bytes = r->v.bytes + 2 * tw * 4;
  {
    // one morphology per request: rounds per iteration and collision points
    // per lane of its template set the time of each of its workgroups
    int32_t ti[6] = {0, 0, 0, 0, 0, 0};
    if (kx_template_info(r->h, 0, ti) >= 0) r->cost = 8 * ti[4] + 4 * ((ti[2] + ti[5] - 1) / std::max(ti[5], 1));
  }
  build_policy(q, r->pol);
  return 0;
}

void kx_server::release(Req *r) {
  for (int s = 0; s < 2; s++) {
    if (r->tape_host[s]) (void)hipHostFree(r->tape_host[s]);
    if (r->tape_dev[s]) (void)hipFree(r->tape_dev[s]);
    r->tape_host[s] = nullptr;
    r->tape_dev[s] = nullptr;
  }
  if (r->h) kx_destroy(r->h);
  r->h = nullptr;
}

void kx_server::finish(Req *r, int status) {
  kx_request_done d{};
  d.tag = r->q.tag;
  d.status = status;
  d.n_creatures = r->ncr;
  d.n_bodies = r->nb;
  d.t_submit = r->t_submit;
  d.t_admit = r->t_admit;
  d.t_done = secs_since(t0);
  d.device_bytes = (int64_t)r->dev_bytes;
  {
    std::lock_guard<std::mutex> g(m);
    done.push_back(d);
    reap.push_back(r);
    n_admitted_live--;               // the slot is free; the memory is until reaped
  }
  n_done++;
  cv_done.notify_all();
  cv_reap.notify_one();
  cv_admit.notify_all();
}

// Frees finished requests off the scheduler's path: freeing device memory
// can wait for the device, which would drain the chunks in flight. A
// request's memory returns to admission when it has been freed here.
void kx_server::reap_loop() {
  if (hipSetDevice(device) != hipSuccess) return;
  for (;;) {
    Req *r = nullptr;
    {
      std::unique_lock<std::mutex> l(m);
      cv_reap.wait(l, [&] { return reap_stop || !reap.empty(); });
      if (reap.empty()) return;
      r = reap.front();
      reap.pop_front();
    }
    release(r);
    delete r;
    n_reaped++;
    cv_admit.notify_all();
  }
}

// Device bytes of a request before its handle exists: the handle's arrays
// (17 words per body, 21 per constraint, 3 per collision point, ~32 per
// creature: spawnKraken makes 1 + arms*segs bodies, arms*(2*segs+1)
// constraints and ring + 2*arms*segs points per creature), two tape slots,
// 1 MiB for the map grid, launch tables and allocation granularity, + 1/8.
static size_t estimate_bytes(const kx_request &q, int K) {
  const size_t ncr = 2 * (size_t)q.pairs;
  const size_t nb = ncr * (size_t)(1 + q.arms * q.segs), nc = ncr * (size_t)(q.arms * (2 * q.segs + 1));
  const size_t np = ncr * (size_t)(q.ring + 2 * q.arms * q.segs);
  const size_t bytes = nb * 68 + nc * 84 + np * 12 + ncr * 128 + 2 * (size_t)K * ncr * KX_TAPE_WORDS * 4 + (1u << 20);
  return bytes + bytes / 8;
}

void kx_server::admit_loop() {
  if (hipSetDevice(device) != hipSuccess) return;
  for (;;) {
    Req *r = nullptr;
    int64_t est = 0;
    {
      std::unique_lock<std::mutex> l(m);
      cv_admit.wait(l, [&] { return stop || (!queued.empty() && n_admitted_live < o.max_active); });
      if (stop) return;
      r = queued.front();
      // free HBM less what the builds in flight are about to allocate
      size_t free_b = 0, total_b = 0;
      est = (int64_t)estimate_bytes(r->q, K);
      const bool mem_ok = hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                          (int64_t)free_b - pending_bytes - est >= o.hbm_reserve;
      if (!mem_ok && n_admitted_live == 0 && reap.empty() && n_reaped.load() == n_done.load()) {
        // nothing resident will ever return memory: the request cannot fit
        queued.pop_front();
        n_admitted_live++;
        l.unlock();
        kx_error(KX_ENOMEM, "request needs more free HBM than the reserve leaves");
        finish(r, KX_ENOMEM);
        continue;
      }
      if (!mem_ok) {
        // wait for a retirement to return memory (or for stop)
        admit_waits++;
        const int seen = n_reaped.load();
        cv_admit.wait_for(l, std::chrono::milliseconds(20), [&] { return stop || n_reaped.load() != seen; });
        continue;
      }
      queued.pop_front();
      n_admitted_live++;
      pending_bytes += est;
    }
    const int rc = build(r);
    {
      std::lock_guard<std::mutex> g(m);
      pending_bytes -= est;
      if (rc >= 0) ready.push_back(r);
    }
    if (rc < 0) {
      r->status = rc;
      finish(r, rc);
      continue;
    }
    cv_sched.notify_all();
  }
}

// ── scheduling ────────────────────────────────────────────────────────
int kx_server::issue(int64_t id, Chunk &c) {
  const int slot = (int)(id & 1);
  c.id = id;
  c.members.clear();
  c.nsteps.clear();
  for (Req *r : active)
    if (r->launched < r->q.steps) {
      c.members.push_back(r);
      c.nsteps.push_back(std::min(K, r->q.steps - r->launched));
    }
  if (c.members.empty()) return 0;
  // policy -> tape rows, tasks of up to 16 pairs
  struct Task { Req *r; int p0, p1, n; };
  std::vector<Task> tasks;
  for (size_t i = 0; i < c.members.size(); i++) {
    Req *r = c.members[i];
    for (int p = 0; p < r->q.pairs; p += 16) tasks.push_back({r, p, std::min(p + 16, r->q.pairs), c.nsteps[i]});
  }
  const double step_s = o.params.h * o.params.substeps;
  pool->run((int)tasks.size(), [&](int t) {
    const Task &k = tasks[t];
    policy_rows(k.r->pol, k.p0, k.p1, k.r->launched, k.n, step_s, k.r->tape_host[slot]);
  });
  // tapes up on the copy stream
  for (size_t i = 0; i < c.members.size(); i++) {
    Req *r = c.members[i];
    const size_t bytes = (size_t)c.nsteps[i] * r->ncr * KX_TAPE_WORDS * 4;
    if (hipMemcpyAsync(r->tape_dev[slot], r->tape_host[slot], bytes, hipMemcpyHostToDevice, copy) != hipSuccess)
      return server_fail(KX_EDEVICE, "tape copy failed");
  }
  if (hipEventRecord(c.copied, copy) != hipSuccess) return server_fail(KX_EDEVICE, "event record failed");
  // batch tables
  const size_t ne = c.members.size();
  size_t nwg = 0, lds = 0;
  for (Req *r : c.members) {
    nwg += (size_t)r->v.P.nwg;
    lds = std::max(lds, kx_persist_lds_bytes(r->v.P));
  }
  if (ne > ent_cap || nwg > map_cap) return server_fail(KX_EINVAL, "batch tables too small");
  KxBatchEnt *E = ent_host[slot];
  int2 *M = map_host[slot];
  KxLaw L{};
  {
    const kx_law &l = o.law;
    L.seed = l.seed;
    L.comp = 1.0 / (l.stiff_min + 0.5 * (l.stiff_max - l.stiff_min));
    L.shoulder_limit = l.shoulder_limit; L.joint_limit = l.joint_limit;
    L.grip_min = l.grip_min; L.grip_max = l.grip_max; L.slip_max = l.slip_max;
    L.cooldown = l.cooldown; L.attach_min = l.attach_min; L.release_below = l.release_below;
    L.stiff_min = l.stiff_min; L.stiff_max = l.stiff_max;
    L.cbase = 0; L.flags = (l.flags & KX_LAW_FIXED_GRIP) | KX_LAW_TAPE;
  }
  // workgroups longest first (cost per step x steps this chunk), so the
  // chunk ends on short workgroups rather than on the largest morphologies
  std::vector<int> order(ne);
  for (size_t i = 0; i < ne; i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (int64_t)c.members[a]->cost * c.nsteps[a] > (int64_t)c.members[b]->cost * c.nsteps[b];
  });
  size_t w = 0;
  for (int i : order)
    for (int b = 0; b < c.members[i]->v.P.nwg; b++) M[w++] = make_int2(i, b);
  for (size_t i = 0; i < ne; i++) {
    Req *r = c.members[i];
    E[i].d = r->v.d;
    E[i].d.tape = r->tape_dev[slot];
    E[i].L = L;
    E[i].P = r->v.P;
    E[i].has_law = 1;
    E[i].step0 = r->launched;
    E[i].nsteps = c.nsteps[i];
    E[i].pad = 0;
  }
  if (hipMemcpyAsync(ent_dev[slot], E, ne * sizeof(KxBatchEnt), hipMemcpyHostToDevice, comp) != hipSuccess ||
      hipMemcpyAsync(map_dev[slot], M, nwg * sizeof(int2), hipMemcpyHostToDevice, comp) != hipSuccess ||
      hipStreamWaitEvent(comp, c.copied, 0) != hipSuccess || hipEventRecord(c.beg, comp) != hipSuccess)
    return server_fail(KX_EDEVICE, "batch staging failed");
  if (kx_persist_launch_batch(1, fold.load() == 1 ? 1 : 0, ent_dev[slot], map_dev[slot], (int)nwg, lds, comp) != 0)
    return server_fail(KX_EDEVICE, "batched launch failed");
  if (hipEventRecord(c.end, comp) != hipSuccess) return server_fail(KX_EDEVICE, "event record failed");
  int64_t cs = 0;
  for (size_t i = 0; i < ne; i++) {
    Req *r = c.members[i];
    if (r->t_admit < 0) r->t_admit = secs_since(t0);
    r->launched += c.nsteps[i];
    if (r->launched >= r->q.steps) r->last_chunk = id;
    cs += (int64_t)c.nsteps[i] * r->ncr;
  }
  launches++;
  creature_steps += cs;
  return 1;
}

void kx_server::retire(Chunk &c) {
  (void)hipEventSynchronize(c.end);
  float ms = 0;
  if (hipEventElapsedTime(&ms, c.beg, c.end) == hipSuccess) {
    std::lock_guard<std::mutex> g(m);
    busy += ms * 1e-3;
  }
  for (Req *r : c.members) {
    if (r->last_chunk != c.id) continue;
    active.erase(std::find(active.begin(), active.end(), r));
    // final body state: px, py, ang, vx, vy, av
    const size_t nb = (size_t)r->nb;
    if (6 * nb > dl_cap) {
      if (dl_host) (void)hipHostFree(dl_host);
      dl_host = nullptr;
      dl_cap = 0;
      if (hipHostMalloc((void **)&dl_host, 6 * nb * 4, hipHostMallocDefault) == hipSuccess) dl_cap = 6 * nb;
    }
    int status = dl_host ? 0 : KX_ENOMEM;
    const float *src[6] = {r->v.d.px, r->v.d.py, r->v.d.ang, r->v.d.vx, r->v.d.vy, r->v.d.av};
    for (int k = 0; k < 6 && !status; k++)
      if (hipMemcpyAsync(dl_host + k * nb, src[k], nb * 4, hipMemcpyDeviceToHost, dl) != hipSuccess) status = KX_EDEVICE;
    if (!status && hipStreamSynchronize(dl) != hipSuccess) status = KX_EDEVICE;
    if (!status) {
      if (r->q.fitness)
        for (int k = 0; k < r->ncr; k++) r->q.fitness[k] = dl_host[(size_t)k * r->bpc] - r->px0[k];
      if (r->q.final_state && r->q.final_cap >= (int64_t)(6 * nb)) std::memcpy(r->q.final_state, dl_host, 6 * nb * 4);
    }
    finish(r, status);
  }
}

void kx_server::sched_loop() {
  if (hipSetDevice(device) != hipSuccess) return;
  Chunk ch[2];
  for (auto &c : ch) {
    (void)hipEventCreate(&c.beg);
    (void)hipEventCreate(&c.end);
    (void)hipEventCreateWithFlags(&c.copied, hipEventDisableTiming);
  }
  std::deque<int> inflight;        // indices into ch, oldest first
  int64_t next_id = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> l(m);
      while (!ready.empty()) {
        active.push_back(ready.front());
        ready.pop_front();
      }
      peak_active = std::max<int64_t>(peak_active.load(), (int64_t)active.size());
      const bool launchable = std::any_of(active.begin(), active.end(), [](Req *r) { return r->launched < r->q.steps; });
      if (!launchable && inflight.empty()) {
        if (stop && ready.empty()) break;
        cv_sched.wait_for(l, std::chrono::milliseconds(5));
        continue;
      }
      if (launchable && inflight.size() < 2) {
        l.unlock();
        const int slot = (int)(next_id & 1);
        const int rc = issue(next_id, ch[slot]);
        if (rc < 0) {
          fprintf(stderr, "kxpbd server: %s\n", kx_last_error());
          for (Req *r : ch[slot].members) {
            active.erase(std::find(active.begin(), active.end(), r));
            finish(r, rc);
          }
          continue;
        }
        if (rc > 0) {
          inflight.push_back(slot);
          next_id++;
        }
        continue;
      }
    }
    if (!inflight.empty()) {
      retire(ch[inflight.front()]);
      inflight.pop_front();
    }
  }
  for (auto &c : ch) {
    (void)hipEventDestroy(c.beg);
    (void)hipEventDestroy(c.end);
    (void)hipEventDestroy(c.copied);
  }
}

extern "C" int kx_server_create(const kx_server_opts *opts, kx_server **out) {
  if (!opts || !out) return server_fail(KX_EINVAL, "null argument");
  *out = nullptr;
  auto *s = new kx_server();
  s->o = *opts;
  s->K = opts->chunk_steps > 0 ? opts->chunk_steps : 50;
  if (s->o.max_active <= 0) s->o.max_active = 64;
  int nt = opts->host_threads > 0 ? opts->host_threads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, 16));
  s->t0 = std::chrono::steady_clock::now();
  if (hipGetDevice(&s->device) != hipSuccess ||
      hipStreamCreateWithFlags(&s->comp, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->dl, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return server_fail(KX_EDEVICE, "stream creation failed");
  }
  s->ent_cap = (size_t)s->o.max_active;
  s->map_cap = 1u << 20;
  for (int k = 0; k < 2; k++) {
    if (hipHostMalloc((void **)&s->ent_host[k], s->ent_cap * sizeof(KxBatchEnt), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->map_host[k], s->map_cap * sizeof(int2), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&s->ent_dev[k], s->ent_cap * sizeof(KxBatchEnt)) != hipSuccess ||
        hipMalloc((void **)&s->map_dev[k], s->map_cap * sizeof(int2)) != hipSuccess) {
      kx_server_destroy(s);
      return server_fail(KX_ENOMEM, "batch table allocation failed");
    }
  }
  s->pool.reset(new Pool(nt - 1));
  for (int k = 0; k < 4; k++) s->th_admit.emplace_back([s] { s->admit_loop(); });
  s->th_reap = std::thread([s] { s->reap_loop(); });
  s->th_sched = std::thread([s] { s->sched_loop(); });
  *out = s;
  return 0;
}

extern "C" int kx_server_submit(kx_server *s, const kx_request *req) {
  if (!s || !req) return server_fail(KX_EINVAL, "null argument");
  const kx_request &q = *req;
  if (q.arms < 1 || q.arms > KX_MAX_ARMS || q.segs < 1 || q.segs > 15 || q.ring < 3 || q.ring > 64)
    return server_fail(KX_EINVAL, "morphology out of range (arms 1-8, segs 1-15, ring 3-64)");
  if (q.pairs < 1 || q.steps < 1) return server_fail(KX_EINVAL, "pairs and steps must be positive");
  if (!std::isfinite(q.sigma) || !std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.angle))
    return server_fail(KX_EINVAL, "non-finite request parameter");
  auto *r = new Req();
  r->q = q;
  r->t_submit = secs_since(s->t0);
  {
    std::lock_guard<std::mutex> g(s->m);
    if (s->stop) {
      delete r;
      return server_fail(KX_EINVAL, "server is stopping");
    }
    s->queued.push_back(r);
  }
  s->cv_admit.notify_all();
  return 0;
}

extern "C" int kx_server_poll(kx_server *s, kx_request_done *out, int32_t cap, int32_t *n, double wait_s) {
  if (!s || !n || (cap > 0 && !out)) return server_fail(KX_EINVAL, "null argument");
  std::unique_lock<std::mutex> l(s->m);
  if (s->done.empty() && wait_s > 0)
    s->cv_done.wait_for(l, std::chrono::duration<double>(wait_s), [&] { return !s->done.empty(); });
  int k = 0;
  while (k < cap && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n = k;
  return 0;
}

extern "C" int kx_server_info_get(kx_server *s, kx_server_info *info) {
  if (!s || !info) return server_fail(KX_EINVAL, "null argument");
  std::lock_guard<std::mutex> g(s->m);
  info->queued = (int32_t)s->queued.size();
  info->active = s->n_admitted_live;
  info->done = s->n_done.load();
  info->launches = s->launches.load();
  info->creature_steps = s->creature_steps.load();
  info->now = secs_since(s->t0);
  info->busy = s->busy;
  info->peak_active = s->peak_active.load();
  info->admit_waits = s->admit_waits.load();
  return 0;
}

extern "C" void kx_server_destroy(kx_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->m);
    s->stop = true;
    for (Req *r : s->queued) delete r;
    s->queued.clear();
  }
  s->cv_admit.notify_all();
  s->cv_sched.notify_all();
  for (auto &t : s->th_admit)
    if (t.joinable()) t.join();
  if (s->th_sched.joinable()) s->th_sched.join();
  {
    std::lock_guard<std::mutex> g(s->m);
    s->reap_stop = true;
  }
  s->cv_reap.notify_all();
  if (s->th_reap.joinable()) s->th_reap.join();
  for (Req *r : s->ready) {
    s->release(r);
    delete r;
  }
  s->pool.reset();
  for (int k = 0; k < 2; k++) {
    if (s->ent_host[k]) (void)hipHostFree(s->ent_host[k]);
    if (s->map_host[k]) (void)hipHostFree(s->map_host[k]);
    if (s->ent_dev[k]) (void)hipFree(s->ent_dev[k]);
    if (s->map_dev[k]) (void)hipFree(s->map_dev[k]);
  }
  if (s->dl_host) (void)hipHostFree(s->dl_host);
  for (hipStream_t st : {s->comp, s->copy, s->dl})
    if (st) (void)hipStreamDestroy(st);
  delete s;
}

extern "C" int64_t kx_request_bytes(const kx_request *req, int32_t chunk_steps) {
  if (!req) return server_fail(KX_EINVAL, "null argument");
  return (int64_t)estimate_bytes(*req, chunk_steps > 0 ? chunk_steps : 50);
}

extern "C" int kx_mem_info(int64_t *free_bytes, int64_t *total_bytes) {
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return server_fail(KX_EDEVICE, "hipMemGetInfo failed");
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return 0;
}

extern "C" int kx_policy_tape(const kx_request *req, int32_t step0, int32_t nsteps, double step_seconds,
                              float *tape) {
  if (!req || (nsteps > 0 && !tape)) return server_fail(KX_EINVAL, "null argument");
  if (req->pairs < 1 || req->arms < 1 || req->arms > KX_MAX_ARMS || step0 < 0 || nsteps < 0)
    return server_fail(KX_EINVAL, "bad request or step range");
  Policy P;
  build_policy(*req, P);
  policy_rows(P, 0, P.pairs, step0, nsteps, step_seconds, tape);
  return 0;
}
