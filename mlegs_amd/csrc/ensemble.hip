// An ensemble of independent vortex runs on one GPU, stepped by replaying
// HIP graphs.
//
// Each member is a full context (its own kit, fields and stream) holding one
// vortex run; members may differ in every parameter (sizes, ell, visc, dt).
// At these sizes one member's AB2/CN step is a chain of some forty small
// launches, so a member alone leaves the GPU idle between them; the
// ensemble keeps the whole chip busy by
//   - capturing each member's step into HIP graphs once (stream capture), so
//     a step is one graph launch instead of forty kernel launches, and
//   - launching the members' graphs on several streams (or as one parent
//     graph), so the steps of different members overlap.
//
// A captured graph holds the device addresses the step used.  The step
// exchanges buffers between fields and the context's scratch (host pointer
// swaps: the Legendre transforms write out of place and swap, the AB2/CN
// update swaps the nonlinear term with its history), so after a step the
// buffers sit in a fixed permutation tau of the slots they started in, and
// a graph captured from one arrangement is wrong for the next.  tau is read
// from one captured step; its order P (the lcm of its cycle lengths) is the
// number of distinct arrangements, the member's "phases".  One graph is
// captured per phase, each starting from that phase's arrangement; a
// member's replays walk through the phases, and the host pointers are set
// to the arrangement the device data is in after each launch.
//
// The launch configuration is tuned on the first call that replays graphs:
// each candidate (one parent graph per phase holding every member's graph
// as a child node; the member graphs launched on 1, 2, 4, 8, 16 or 32
// streams from one host thread; or the same with one host thread per
// stream, as a graph launch costs host time per node) runs 1 + 2 rounds and
// the fastest is kept; then the winner is tried with graphs of four steps
// each (fewer, larger launches) and kept if faster.  The tuning rounds are ordinary steps of the run: every
// configuration issues the same kernels in the same order per member, so a
// member's results do not depend on the configuration.
//
// The reference runs one kit per process (mlegs_spectfm_init.f90:15-16); this
// has no counterpart there.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "mlx_internal.h"

using namespace mlx;

namespace {

constexpr int kMaxPhases = 64;

typedef std::vector<void *> Layout;

// every device buffer a member's step may exchange, slot by slot: the run's
// fields, the context's temporaries and its scratch buffers
Layout layout_of(mlx_ctx *c, mlx_vortex *v) {
  Layout l;
  for (mlx_field *f : v->f) {
    l.push_back(f->e);
    l.push_back(f->sc);
  }
  for (mlx_field *f : c->pool) {
    l.push_back(f ? (void *)f->e : nullptr);
    l.push_back(f ? (void *)f->sc : nullptr);
  }
  l.push_back(c->scr[0]);
  l.push_back(c->scr[1]);
  return l;
}

void set_layout(mlx_ctx *c, mlx_vortex *v, const Layout &l) {
  size_t i = 0;
  for (mlx_field *f : v->f) {
    f->e = (double2 *)l[i++];
    f->sc = (double *)l[i++];
  }
  for (mlx_field *f : c->pool) {
    if (f) {
      f->e = (double2 *)l[i];
      f->sc = (double *)l[i + 1];
    }
    i += 2;
  }
  c->scr[0] = (double2 *)l[i++];
  c->scr[1] = (double2 *)l[i++];
}

struct Member {
  mlx_ctx *c = nullptr;
  mlx_vortex *v = nullptr;
  bool warm = false;                 // stepped once eagerly: every temporary exists
  int phases = 0;                    // order of the step's slot permutation (0: not captured)
  int steps = 0;                     // steps per graph
  int phase = 0;                     // the arrangement the device data is in now
  std::vector<Layout> lay;           // arrangement of each phase
  std::vector<hipGraph_t> g;         // graph of `steps` steps starting from each phase
  std::vector<hipGraphExec_t> x;
};

void step_eager(Member &m, int n) {
  if (n <= 0) return;
  int rc = mlx_vortex_step(m.c, m.v, n, m.c->p.dt, 0);
  if (rc != MLX_OK) throw MlxError{rc, m.c->err};
}

void drop_graphs(Member &m) {
  for (hipGraphExec_t e : m.x)
    if (e) (void)hipGraphExecDestroy(e);
  for (hipGraph_t g : m.g)
    if (g) (void)hipGraphDestroy(g);
  m.x.clear();
  m.g.clear();
  m.lay.clear();
  m.phases = 0;
  m.steps = 0;
  m.phase = 0;
}

// Capture n steps from the current host arrangement; the host arrangement
// is put back afterwards (capturing issues nothing to the device).  Returns
// the graph and the arrangement the n steps end in.
hipGraph_t capture_steps(Member &m, int n, Layout *end) {
  mlx_ctx *c = m.c;
  const Layout start = layout_of(c, m.v);
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = mlx_vortex_step(c, m.v, n, c->p.dt, 0);
  hipError_t er = hipStreamEndCapture(c->stream, &g);
  *end = layout_of(c, m.v);
  set_layout(c, m.v, start);
  if (rc != MLX_OK) {
    if (g) (void)hipGraphDestroy(g);
    throw MlxError{rc, "ensemble: capture of a member step failed: " + c->err};
  }
  if (er != hipSuccess) MLX_THROW(MLX_E_HIP, std::string("ensemble: stream capture failed: ") + hipGetErrorString(er));
  return g;
}

// Graphs of `steps` steps for every phase of the member's step permutation.
void capture(Member &m, int steps) {
  drop_graphs(m);
  const Layout l0 = layout_of(m.c, m.v);
  Layout l1;
  hipGraph_t g1 = capture_steps(m, 1, &l1);
  (void)hipGraphDestroy(g1);
  // tau: slot i holds after a step what slot tau[i] held before it
  const size_t n = l0.size();
  std::vector<size_t> tau(n);
  for (size_t i = 0; i < n; ++i) {
    size_t j = 0;
    while (j < n && l0[j] != l1[i]) ++j;
    if (j == n) MLX_THROW(MLX_E_ARG, "ensemble: the step left a buffer outside its fields");
    tau[i] = j;
  }
  long order = 1;
  std::vector<bool> seen(n, false);
  for (size_t i = 0; i < n; ++i) {
    if (seen[i]) continue;
    long len = 0;
    for (size_t j = i; !seen[j]; j = tau[j]) {
      seen[j] = true;
      ++len;
    }
    order = std::lcm(order, len);
    if (order > kMaxPhases) MLX_THROW(MLX_E_ARG, "ensemble: the step's buffer permutation has too many phases");
  }
  m.phases = (int)order;
  m.steps = steps;
  m.lay.resize(order);
  m.lay[0] = l0;
  for (long p = 1; p < order; ++p) {
    m.lay[p].resize(n);
    for (size_t i = 0; i < n; ++i) m.lay[p][i] = m.lay[p - 1][tau[i]];
  }
  m.g.assign(order, nullptr);
  m.x.assign(order, nullptr);
  for (long p = 0; p < order; ++p) {
    set_layout(m.c, m.v, m.lay[p]);
    Layout end;
    m.g[p] = capture_steps(m, steps, &end);
    if (end != m.lay[(p + steps) % order]) {
      set_layout(m.c, m.v, l0);
      MLX_THROW(MLX_E_ARG, "ensemble: a captured step did not follow the step's buffer permutation");
    }
    HIPCHK(hipGraphInstantiate(&m.x[p], m.g[p], nullptr, nullptr, 0));
  }
  set_layout(m.c, m.v, l0);
  m.phase = 0;
}

// the phase of the member's current host arrangement, or -1
int find_phase(Member &m) {
  if (!m.phases) return -1;
  const Layout now = layout_of(m.c, m.v);
  for (int p = 0; p < m.phases; ++p)
    if (now == m.lay[p]) return p;
  return -1;
}

}  // namespace

struct mlx_ensemble {
  std::vector<Member> m;
  int device = 0;
  std::string err;
  // launch configuration: mode 0 parent graphs, mode 1 member graphs on
  // `streams` streams, mode 2 the same with one host thread per stream
  int mode = -1, streams = 0, steps = 1;
  std::vector<hipStream_t> s;        // launch streams (default priority)
  std::vector<hipGraph_t> parent;    // per phase (all members in the same phase)
  std::vector<hipGraphExec_t> parent_x;
  hipEvent_t ev = nullptr;
  std::vector<hipEvent_t> evs;
  double tune_ms[3][6][2] = {};      // [mode][streams index][steps 1, 4] ms per step of all members
};

namespace {

int guarded_e(mlx_ensemble *e, const std::function<void()> &fn) {
  if (!e) return MLX_E_ARG;
  try {
    HIPCHK(hipSetDevice(e->device));
    fn();
    return MLX_OK;
  } catch (const MlxError &x) {
    e->err = x.msg;
    return x.code;
  } catch (const std::exception &x) {
    e->err = x.what();
    return MLX_E_ARG;
  }
}

void drop_parent(mlx_ensemble *e) {
  for (hipGraphExec_t x : e->parent_x)
    if (x) (void)hipGraphExecDestroy(x);
  for (hipGraph_t g : e->parent)
    if (g) (void)hipGraphDestroy(g);
  e->parent_x.clear();
  e->parent.clear();
}

void ensure_streams(mlx_ensemble *e, int n) {
  while ((int)e->s.size() < n) {
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    e->s.push_back(st);
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    e->evs.push_back(ev);
  }
  if (!e->ev) HIPCHK(hipEventCreateWithFlags(&e->ev, hipEventDisableTiming));
}

// launch streams wait for what was issued on the member streams
void fence_members_to(mlx_ensemble *e, int nstreams) {
  for (Member &m : e->m) {
    HIPCHK(hipEventRecord(e->ev, m.c->stream));
    for (int i = 0; i < nstreams; ++i) HIPCHK(hipStreamWaitEvent(e->s[i], e->ev, 0));
  }
}
// member streams wait for the launch streams, before eager work follows
void fence_streams_to_members(mlx_ensemble *e, int nstreams) {
  for (int i = 0; i < nstreams; ++i) HIPCHK(hipEventRecord(e->evs[i], e->s[i]));
  for (Member &m : e->m)
    for (int i = 0; i < nstreams; ++i) HIPCHK(hipStreamWaitEvent(m.c->stream, e->evs[i], 0));
}

// every member captured with `steps` steps per graph, in a known phase
bool ready(mlx_ensemble *e, int steps) {
  for (Member &m : e->m) {
    if (m.steps != steps) return false;
    const int p = find_phase(m);
    if (p < 0) return false;
    m.phase = p;
  }
  return true;
}

void capture_all(mlx_ensemble *e, int steps) {
  drop_parent(e);
  for (Member &m : e->m) capture(m, steps);
  e->steps = steps;
}

// parent graphs need every member with the same phase count and phase
bool parent_possible(mlx_ensemble *e) {
  for (Member &m : e->m)
    if (m.phases != e->m[0].phases || m.phase != e->m[0].phase) return false;
  return true;
}

void build_parents(mlx_ensemble *e) {
  drop_parent(e);
  const int np = e->m[0].phases;
  e->parent.assign(np, nullptr);
  e->parent_x.assign(np, nullptr);
  for (int p = 0; p < np; ++p) {
    HIPCHK(hipGraphCreate(&e->parent[p], 0));
    for (Member &m : e->m) {
      hipGraphNode_t node;
      HIPCHK(hipGraphAddChildGraphNode(&node, e->parent[p], nullptr, 0, m.g[p]));
    }
    HIPCHK(hipGraphInstantiate(&e->parent_x[p], e->parent[p], nullptr, nullptr, 0));
  }
}

void advance(Member &m) {
  m.phase = (m.phase + m.steps) % m.phases;
  set_layout(m.c, m.v, m.lay[m.phase]);
}

// one round: every member advances by e->steps steps
void replay(mlx_ensemble *e, int mode, int nstreams) {
  if (mode == 0) {
    HIPCHK(hipGraphLaunch(e->parent_x[e->m[0].phase], e->s[0]));
    for (Member &m : e->m) advance(m);
  } else {
    for (size_t i = 0; i < e->m.size(); ++i) {
      Member &m = e->m[i];
      HIPCHK(hipGraphLaunch(m.x[m.phase], e->s[i % nstreams]));
      advance(m);
    }
  }
}

// `rounds` rounds of mode 2: stream t is fed by its own host thread, which
// launches the graphs of members t, t + nstreams, ... round after round
// (members are independent, so each thread touches only its own members)
void replay_threaded(mlx_ensemble *e, int nstreams, int rounds) {
  std::vector<std::thread> th;
  std::vector<std::string> errs(nstreams);
  for (int t = 0; t < nstreams; ++t) {
    th.emplace_back([e, t, nstreams, rounds, &errs] {
      if (hipSetDevice(e->device) != hipSuccess) {
        errs[t] = "hipSetDevice failed";
        return;
      }
      for (int r = 0; r < rounds; ++r)
        for (size_t i = t; i < e->m.size(); i += nstreams) {
          Member &m = e->m[i];
          hipError_t er = hipGraphLaunch(m.x[m.phase], e->s[t]);
          if (er != hipSuccess) {
            errs[t] = std::string("hipGraphLaunch: ") + hipGetErrorString(er);
            return;
          }
          advance(m);
        }
    });
  }
  for (std::thread &x : th) x.join();
  for (const std::string &s : errs)
    if (!s.empty()) MLX_THROW(MLX_E_HIP, "ensemble: " + s);
}

void sync_streams(mlx_ensemble *e, int nstreams) {
  for (int i = 0; i < nstreams; ++i) HIPCHK(hipStreamSynchronize(e->s[i]));
}

const int kStreamChoices[6] = {1, 2, 4, 8, 16, 32};

// 1 + count rounds with the given configuration; ms per timed round
double run_rounds(mlx_ensemble *e, int mode, int nstreams, int count) {
  if (mode == 0 && e->parent_x.empty()) build_parents(e);
  fence_members_to(e, nstreams);
  replay(e, mode == 2 ? 1 : mode, nstreams);  // untimed: the first launch of an executable graph uploads it
  sync_streams(e, nstreams);
  auto t0 = std::chrono::steady_clock::now();
  if (mode == 2)
    replay_threaded(e, nstreams, count);
  else
    for (int r = 0; r < count; ++r) replay(e, mode, nstreams);
  sync_streams(e, nstreams);
  auto t1 = std::chrono::steady_clock::now();
  fence_streams_to_members(e, nstreams);
  return std::chrono::duration<double, std::milli>(t1 - t0).count() / std::max(count, 1);
}

}  // namespace

extern "C" {

int mlx_ensemble_create(const mlx_params *p, int32_t n, int device, int32_t prec, mlx_ensemble **out) {
  if (!p || !out || n <= 0) return MLX_E_ARG;
  *out = nullptr;
  mlx_ensemble *e = new mlx_ensemble();
  e->device = device;
  int rc = guarded_e(e, [&] {
    e->m.resize(n);
    for (int i = 0; i < n; ++i) {
      int r = mlx_ctx_create(&p[i], device, nullptr, &e->m[i].c);
      if (r != MLX_OK) MLX_THROW(r, "ensemble: member " + std::to_string(i) + ": context creation failed");
      r = mlx_kit_init(e->m[i].c);
      if (r != MLX_OK) MLX_THROW(r, "ensemble: member " + std::to_string(i) + ": " + e->m[i].c->err);
      r = mlx_set_precision(e->m[i].c, prec);
      if (r != MLX_OK) MLX_THROW(r, "ensemble: member " + std::to_string(i) + ": " + e->m[i].c->err);
    }
  });
  if (rc != MLX_OK) {
    for (Member &m : e->m)
      if (m.c) mlx_ctx_destroy(m.c);
    delete e;
    return rc;
  }
  *out = e;
  return MLX_OK;
}

int mlx_ensemble_destroy(mlx_ensemble *e) {
  if (!e) return MLX_OK;
  (void)hipSetDevice(e->device);
  for (hipStream_t st : e->s) (void)hipStreamSynchronize(st);
  drop_parent(e);
  for (Member &m : e->m) {
    if (m.c) (void)hipStreamSynchronize(m.c->stream);
    drop_graphs(m);
    if (m.v) mlx_vortex_destroy(m.c, m.v);
    if (m.c) mlx_ctx_destroy(m.c);
  }
  for (hipStream_t st : e->s) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : e->evs) (void)hipEventDestroy(ev);
  if (e->ev) (void)hipEventDestroy(e->ev);
  delete e;
  return MLX_OK;
}

const char *mlx_ensemble_last_error(mlx_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

int mlx_ensemble_ctx(mlx_ensemble *e, int32_t i, mlx_ctx **ctx) {
  return guarded_e(e, [&] {
    if (i < 0 || i >= (int)e->m.size() || !ctx) MLX_THROW(MLX_E_ARG, "ensemble: member index out of range");
    *ctx = e->m[i].c;
  });
}

int mlx_ensemble_set(mlx_ensemble *e, int32_t i, const mlx_field *psi, const mlx_field *chi, const mlx_field *uz) {
  return guarded_e(e, [&] {
    if (i < 0 || i >= (int)e->m.size()) MLX_THROW(MLX_E_ARG, "ensemble: member index out of range");
    Member &m = e->m[i];
    drop_parent(e);
    drop_graphs(m);
    if (m.v) {
      mlx_vortex_destroy(m.c, m.v);
      m.v = nullptr;
    }
    m.warm = false;
    int r = mlx_vortex_create(m.c, psi, chi, uz, &m.v);
    if (r != MLX_OK) throw MlxError{r, m.c->err};
  });
}

int mlx_ensemble_vortex(mlx_ensemble *e, int32_t i, mlx_vortex **v) {
  return guarded_e(e, [&] {
    if (i < 0 || i >= (int)e->m.size() || !v) MLX_THROW(MLX_E_ARG, "ensemble: member index out of range");
    if (!e->m[i].v) MLX_THROW(MLX_E_ARG, "ensemble: member has no state (mlx_ensemble_set)");
    for (hipStream_t st : e->s) HIPCHK(hipStreamSynchronize(st));
    *v = e->m[i].v;
  });
}

int mlx_ensemble_bootstrap(mlx_ensemble *e) {
  return guarded_e(e, [&] {
    for (Member &m : e->m)
      if (!m.v) MLX_THROW(MLX_E_ARG, "ensemble: every member needs a state (mlx_ensemble_set)");
    drop_parent(e);
    for (Member &m : e->m) {
      drop_graphs(m);
      int r = mlx_vortex_bootstrap(m.c, m.v);
      if (r != MLX_OK) throw MlxError{r, m.c->err};
    }
  });
}

int mlx_ensemble_step(mlx_ensemble *e, int32_t nsteps, int32_t mode, int32_t nstreams, int32_t reps) {
  return guarded_e(e, [&] {
    if (nsteps < 0) MLX_THROW(MLX_E_ARG, "ensemble: nsteps must not be negative");
    if (mode < -1 || mode > 2)
      MLX_THROW(MLX_E_ARG, "ensemble: mode is -1 (tune), 0 (parent graph), 1 (streams) or 2 (threaded streams)");
    if (mode >= 1 && (nstreams < 1 || nstreams > 64)) MLX_THROW(MLX_E_ARG, "ensemble: 1 to 64 streams");
    if (mode >= 0 && reps != 1 && reps != 4) MLX_THROW(MLX_E_ARG, "ensemble: reps is 1 or 4");
    for (Member &m : e->m)
      if (!m.v) MLX_THROW(MLX_E_ARG, "ensemble: every member needs a state (mlx_ensemble_set)");
    int left = nsteps;
    // a member's first step is eager: it allocates every temporary the step
    // uses, which a capture must not do
    bool cold = false;
    for (Member &m : e->m) cold = cold || !m.warm;
    if (cold && left > 0) {
      for (Member &m : e->m) {
        step_eager(m, 1);
        m.warm = true;
      }
      --left;
    }
    if (left == 0) return;
    const int want = mode >= 0 ? reps : (e->mode >= 0 ? e->steps : 1);
    if (!ready(e, want)) capture_all(e, want);
    if (mode >= 0) {
      if (mode == 0 && !parent_possible(e)) MLX_THROW(MLX_E_ARG, "ensemble: members' step phases differ; use mode 1");
      e->mode = mode;
      e->streams = mode == 0 ? 1 : nstreams;
      drop_parent(e);
    } else if (e->mode < 0) {
      // tune: every candidate replays 1 + 2 rounds (ordinary steps of the run)
      ensure_streams(e, 32);
      double best = 1e300;
      int bm = 1, bs = 8;
      for (int md = 0; md < 3; ++md) {
        if (md == 0 && !parent_possible(e)) continue;
        for (int si = 0; si < 6; ++si) {
          if (md == 0 && si > 0) break;
          if (left < 3 * e->steps) break;
          const int ns = md == 0 ? 1 : kStreamChoices[si];
          const double per_step = run_rounds(e, md, ns, 2) / e->steps;
          left -= 3 * e->steps;
          e->tune_ms[md][si][0] = per_step;
          if (per_step < best) {
            best = per_step;
            bm = md;
            bs = ns;
          }
        }
      }
      e->mode = bm;
      e->streams = bs;
      // graphs of four steps each, with the best launch configuration
      if (left >= 2 * 3 * 4) {
        capture_all(e, 4);
        if (bm >= 1 || parent_possible(e)) {
          const double per_step = run_rounds(e, bm, bs, 2) / 4;
          left -= 3 * 4;
          int si = 0;
          while (kStreamChoices[si] != bs) ++si;
          e->tune_ms[bm][si][1] = per_step;
          if (per_step >= best) capture_all(e, 1);
        } else {
          capture_all(e, 1);
        }
      }
    }
    ensure_streams(e, std::max(e->streams, 1));
    if (e->mode == 0 && !parent_possible(e)) e->mode = 1;
    const int rounds = left / e->steps;
    if (rounds > 0) {
      if (e->mode == 0 && e->parent_x.empty()) build_parents(e);
      fence_members_to(e, e->streams);
      if (e->mode == 2)
        replay_threaded(e, e->streams, rounds);
      else
        for (int r = 0; r < rounds; ++r) replay(e, e->mode, e->streams);
      fence_streams_to_members(e, e->streams);
      left -= rounds * e->steps;
    }
    for (Member &m : e->m) step_eager(m, left);
  });
}

int mlx_ensemble_sync(mlx_ensemble *e) {
  return guarded_e(e, [&] {
    for (hipStream_t st : e->s) HIPCHK(hipStreamSynchronize(st));
    for (Member &m : e->m) HIPCHK(hipStreamSynchronize(m.c->stream));
  });
}

// the first member whose state holds a non-finite value, or -1
int mlx_ensemble_check(mlx_ensemble *e, int32_t *bad) {
  return guarded_e(e, [&] {
    if (!bad) MLX_THROW(MLX_E_ARG, "ensemble: null output");
    *bad = -1;
    for (hipStream_t st : e->s) HIPCHK(hipStreamSynchronize(st));
    for (size_t i = 0; i < e->m.size(); ++i) {
      Member &m = e->m[i];
      if (!m.v) continue;
      int h = 0;
      HIPCHK(hipMemcpyAsync(&h, m.v->flag, sizeof(int), hipMemcpyDeviceToHost, m.c->stream));
      HIPCHK(hipStreamSynchronize(m.c->stream));
      if (h && *bad < 0) *bad = (int32_t)i;
    }
  });
}

int mlx_ensemble_info(mlx_ensemble *e, int32_t *info, double *tune) {
  return guarded_e(e, [&] {
    if (!info) MLX_THROW(MLX_E_ARG, "ensemble: null output");
    info[0] = e->mode;
    info[1] = e->streams;
    info[2] = e->steps;
    info[3] = e->m.empty() ? 0 : e->m[0].phases;
    info[4] = (int32_t)e->m.size();
    if (tune) std::memcpy(tune, e->tune_ms, sizeof(e->tune_ms));
  });
}

}  // extern "C"
