// Datadeps executor: a static task DAG run on HIP streams / hipGraphs.
//
// Semantics follow Dagger's datadeps regions (src/datadeps/queue.jl:44-54,
// dependency rules of get_write_deps!/get_read_deps! in
// src/datadeps/aliasing.jl:740-773): tasks are recorded in submission
// order; an In access waits for the last earlier overlapping writer, an
// Out/InOut access waits for every earlier overlapping reader and writer,
// In/In never synchronize.  Overlap is decided on byte spans of the operand
// footprints (column-major rows/cols/ld), which for whole tiles is Dagger's
// ContiguousAliasing of the tile array (src/memory-spaces.jl:466-470).
//
// Planning (done once, reused by every run):
//  1. edges from the rules above;
//  2. ASAP level (longest path from a source) and ALAP level (latest level
//     that does not lengthen the DAG); slack = ALAP - ASAP;
//  3. groups: tasks with the same ASAP level, the same kernel shape and the
//     same criticality (slack <= lookahead) become one batched launch -- a
//     level holds no edges, so its tasks are independent;
//  4. critical groups go on stream 0, others on streams 1..S-1; a group
//     waits (hipStreamWaitEvent) on the newest predecessor group of every
//     other stream; groups are issued in level order;
//  5. optionally the whole issue sequence is captured into a hipGraph.
// For the tiled Cholesky this yields the classic look-ahead schedule: the
// diagonal POTRF, the panel TRSM and the next panel's update run on the
// critical stream while the rest of the trailing update runs beside them.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <type_traits>
#include <vector>
#include "dm_internal.h"
#include "dag.h"
#include "complex_ops.h"

namespace dm {

namespace {

size_t esize(int dtype) {
  return dtype == DM_F32 ? 4 : dtype == DM_C64 ? 16 : 8;  // DM_F64, DM_C32: 8
}

// Footprint of a column-major rows x cols array with leading dimension ld.
void mat_span(const void* p, long rows, long cols, long ld, size_t es, uintptr_t& lo,
              uintptr_t& hi) {
  lo = (uintptr_t)p;
  if (rows <= 0 || cols <= 0) { hi = lo; return; }
  hi = lo + (size_t)((cols - 1) * ld + rows) * es;
}

}  // namespace

int Dag::add(const dm_task& t) {
  if (planned) return arg_fail(1, "dag already planned");
  if (t.op < DM_OP_POTRF || t.op > DM_OP_LASWP) return arg_fail(2, "unknown op");
  if (t.dtype != DM_F64 && t.dtype != DM_F32 && t.dtype != DM_C64 && t.dtype != DM_C32)
    return arg_fail(2, "unknown dtype");
  if (is_cplx(t.dtype) && t.op != DM_OP_POTRF && t.op != DM_OP_TRSM && t.op != DM_OP_SYRK &&
      t.op != DM_OP_GEMM)
    return arg_fail(2, "complex tasks are POTRF, TRSM, SYRK (herk) and GEMM");
  TaskRec r;
  r.t = t;
  const size_t es = esize(t.dtype);
  auto push = [&](const void* p, long rows, long cols, long ld, int mode) {
    if (!p || mode == 0) return;
    Access a;
    mat_span(p, rows, cols, ld, es, a.lo, a.hi);
    a.mode = mode;
    if (a.hi > a.lo) r.acc.push_back(a);
  };
  const bool ta = is_trans(t.transa), tb = is_trans(t.transb);
  switch (t.op) {
    case DM_OP_GEMM:
      push(t.a, ta ? t.k : t.m, ta ? t.m : t.k, t.lda, t.access[0]);
      push(t.b, tb ? t.n : t.k, tb ? t.k : t.n, t.ldb, t.access[1]);
      push(t.c, t.m, t.n, t.ldc, t.access[2]);
      r.cost = 2.0 * t.m * t.n * t.k;
      break;
    case DM_OP_SYRK:
      push(t.a, ta ? t.k : t.n, ta ? t.n : t.k, t.lda, t.access[0]);
      push(t.c, t.n, t.n, t.ldc, t.access[2]);
      r.cost = 1.0 * t.n * t.n * t.k;
      break;
    case DM_OP_TRSM: {
      const long ka = is_left(t.side) ? t.m : t.n;
      push(t.a, ka, ka, t.lda, t.access[0]);
      push(t.c, t.m, t.n, t.ldc, t.access[2]);
      r.cost = 1.0 * t.m * t.n * ka;
      break;
    }
    case DM_OP_POTRF:
      push(t.c, t.n, t.n, t.ldc, t.access[2] ? t.access[2] : DM_INOUT);
      if (t.info) push(t.info, 1, 1, 1, DM_OUT);
      r.cost = 1.0 * t.n * t.n * t.n / 3.0;
      break;
    case DM_OP_GEMV: {
      // x in slot b (stride ldb), y in slot c (stride ldc)
      const long lx = ta ? t.m : t.n, ly = ta ? t.n : t.m;
      push(t.a, t.m, t.n, t.lda, t.access[0]);
      push(t.b, 1, lx, t.ldb, t.access[1]);
      push(t.c, 1, ly, t.ldc, t.access[2]);
      r.cost = 2.0 * t.m * t.n;
      break;
    }
    case DM_OP_TRSV:
      push(t.a, t.n, t.n, t.lda, t.access[0]);
      push(t.c, 1, t.n, t.ldc, t.access[2]);
      r.cost = 1.0 * t.n * t.n;
      break;
    case DM_OP_GETRF:
    case DM_OP_LASWP: {
      if (t.m < 0 || t.n < 0 || t.k <= 0) return arg_fail(2, "getrf/laswp: bad panel shape");
      // the block-column region: m x n elements, tiles packed contiguously
      push(t.c, (long)t.m * t.n, 1, (long)t.m * t.n, t.access[2] ? t.access[2] : DM_INOUT);
      const long npiv = t.op == DM_OP_GETRF ? std::min(t.m, t.n) : t.lda;
      auto push_int = [&](const void* p, long count, int mode) {
        if (!p || mode == 0 || count <= 0) return;
        Access a;
        mat_span(p, count, 1, count, 4, a.lo, a.hi);
        a.mode = mode;
        r.acc.push_back(a);
      };
      push_int(t.b, npiv, t.access[1] ? t.access[1] : (t.op == DM_OP_GETRF ? DM_OUT : DM_IN));
      if (t.op == DM_OP_GETRF) push_int(t.info, 1, DM_INOUT);
      r.cost = t.op == DM_OP_GETRF ? 1.0 * t.m * t.n * t.n : 1.0 * t.m * t.n;
      break;
    }
  }
  const int id = (int)tasks.size();
  // ---- dependencies (datadeps rules on overlapping spans)
  std::set<int> deps;
  for (const Access& a : r.acc) {
    const bool writes = (a.mode & DM_OUT) != 0;
    auto it = regions_by_lo.lower_bound(a.lo >= max_len ? a.lo - max_len : 0);
    for (; it != regions_by_lo.end() && it->first < a.hi; ++it) {
      const Region& g = regions[it->second];
      if (g.hi <= a.lo) continue;
      if (g.last_writer >= 0) deps.insert(g.last_writer);
      if (writes)
        for (int rd : g.readers) deps.insert(rd);
    }
  }
  deps.erase(id);
  r.preds.assign(deps.begin(), deps.end());
  // ---- update region state
  for (const Access& a : r.acc) {
    int gi = -1;
    auto range = regions_by_lo.equal_range(a.lo);
    for (auto it = range.first; it != range.second; ++it)
      if (regions[it->second].hi == a.hi) { gi = it->second; break; }
    if (gi < 0) {
      gi = (int)regions.size();
      Region g;
      g.lo = a.lo;
      g.hi = a.hi;
      regions.push_back(g);
      regions_by_lo.emplace(a.lo, gi);
      max_len = std::max<uintptr_t>(max_len, a.hi - a.lo);
    }
    Region& g = regions[gi];
    if (a.mode & DM_OUT) {
      g.last_writer = id;
      g.readers.clear();
    } else if (g.readers.empty() || g.readers.back() != id) {
      g.readers.push_back(id);
    }
  }
  for (int p : r.preds) tasks[p].succs.push_back(id);
  tasks.push_back(std::move(r));
  return 0;
}

// Group key: kernel shape that can share one batched launch.
static long shape_key(const dm_task& t) {
  long key = (long)t.dtype << 20;
  switch (t.op) {
    case DM_OP_GEMM:
    case DM_OP_SYRK: {
      bool ta, tb;
      if (t.op == DM_OP_GEMM) { ta = is_trans(t.transa); tb = is_trans(t.transb); }
      else { ta = is_trans(t.transa); tb = !ta; }
      return key | (1L << 8) | (ta ? 2 : 0) | (tb ? 1 : 0);
    }
    case DM_OP_TRSM: {
      TrsmShape sh = trsm_shape(t.side, t.uplo, t.transa, t.diag);
      return key | (2L << 8) | (sh.lower ? 8 : 0) | (sh.ttrans ? 4 : 0) | (sh.side_l ? 2 : 0) |
             (sh.unit ? 1 : 0);
    }
    default:
      return key | ((long)t.op << 8);
  }
}

int Dag::plan(const dm_plan_opts& o) {
  if (planned) return arg_fail(1, "dag already planned");
  opts = o;
  if (opts.num_streams < 1) opts.num_streams = 1;
  if (opts.num_streams > 8) opts.num_streams = 8;
  const int n = (int)tasks.size();
  // Wave priority goes to the critical stream only in plans with
  // factorization tasks, whose panel chain (POTRF/GETRF, then the panel
  // solves) bounds the wall.  In a blocked substitution (GEMM-only sweeps)
  // the bulk launches set each step's period and share the CUs with the
  // critical launches, so they get it instead: the ldiv! sweeps at
  // N = 32768 with 64 right-hand sides took 6.75 ms with the critical
  // launches raised, 6.55 ms with neither, 6.21 ms with the bulk raised.
  crit_prio = false;
  for (const TaskRec& r : tasks)
    if (r.t.op == DM_OP_POTRF || r.t.op == DM_OP_GETRF) crit_prio = true;
  // ---- levels
  int maxlev = -1;
  for (int i = 0; i < n; ++i) {  // submission order is a topological order
    int lv = 0;
    for (int p : tasks[i].preds) lv = std::max(lv, tasks[p].level + 1);
    tasks[i].level = lv;
    maxlev = std::max(maxlev, lv);
  }
  for (int i = n - 1; i >= 0; --i) {
    int al = maxlev;
    for (int s : tasks[i].succs) al = std::min(al, tasks[s].alap - 1);
    tasks[i].alap = al;
  }
  num_levels = maxlev + 1;
  // ---- groups
  std::map<std::tuple<int, int, long, int>, int> gid;  // (level, crit, shape, serial) -> group
  for (int i = 0; i < n; ++i) {
    TaskRec& r = tasks[i];
    const int slack = r.alap - r.level;
    const int crit = (opts.lookahead > 0 && slack <= opts.lookahead) ? 0 : 1;
    // complex tasks launch one by one (complex.hip), real ones batch
    const bool batchable = opts.batch && !is_cplx(r.t.dtype) &&
                           (r.t.op == DM_OP_GEMM || r.t.op == DM_OP_SYRK || r.t.op == DM_OP_TRSM);
    const auto key = std::make_tuple(r.level, crit, shape_key(r.t), batchable ? -1 : i);
    auto it = gid.find(key);
    int g;
    if (it == gid.end()) {
      g = (int)groups.size();
      gid.emplace(key, g);
      Group G;
      G.level = r.level;
      G.crit = crit;
      G.op = r.t.op == DM_OP_SYRK ? DM_OP_GEMM : r.t.op;
      G.dtype = r.t.dtype;
      groups.push_back(G);
    } else {
      g = it->second;
    }
    groups[g].tasks.push_back(i);
    groups[g].flops += r.cost;
    r.group = g;
  }
  // issue order: by level, critical first (std::map order already is)
  order.clear();
  for (auto& kv : gid) order.push_back(kv.second);
  // ---- streams
  int rr = 0;
  for (int g : order) {
    Group& G = groups[g];
    if (opts.num_streams == 1 || (opts.lookahead > 0 && G.crit == 0)) G.stream = 0;
    else if (opts.lookahead > 0) G.stream = 1 + (rr++ % (opts.num_streams - 1));
    else G.stream = rr++ % opts.num_streams;
  }
  // ---- cross-stream waits: newest predecessor group per other stream
  std::vector<int> issue_pos(groups.size());
  for (size_t i = 0; i < order.size(); ++i) issue_pos[order[i]] = (int)i;
  for (int g : order) {
    Group& G = groups[g];
    std::vector<int> newest(opts.num_streams, -1);
    for (int t : G.tasks)
      for (int p : tasks[t].preds) {
        const int pg = tasks[p].group;
        const Group& P = groups[pg];
        if (P.stream == G.stream) continue;
        if (newest[P.stream] < 0 || issue_pos[pg] > issue_pos[newest[P.stream]])
          newest[P.stream] = pg;
      }
    for (int sidx = 0; sidx < opts.num_streams; ++sidx)
      if (newest[sidx] >= 0) G.waits.push_back(newest[sidx]);
  }
  int rc = build_device_tables();
  if (rc) return rc;
  // Graph capture is limited to plans with at most two worker streams (a
  // capture of the three-stream look-ahead-2 schedule sent the runtime's
  // graph walk into unbounded recursion, a stack overflow in libamdhip64)
  // and to plans small enough for launch overhead to matter (larger plans
  // issue eagerly).
  double total_flops = 0;
  for (const Group& G : groups) total_flops += G.flops;
  const bool want_graph =
      opts.use_graph && opts.num_streams <= 2 && total_flops <= kGraphMaxFlops;
  rc = create_streams(!want_graph);
  if (rc) return rc;
  planned = true;
  if (want_graph) {
    rc = capture();
    if (rc) return rc;
  }
  return 0;
}

int Dag::create_streams(bool eager) {
  // An eagerly issued plan runs on the executor streams of its caller's
  // stream (shared by all plans issued from it); a plan captured into a
  // hipGraph owns its streams and the origin stream the capture runs on.
  if (!eager) {
    streams.resize(opts.num_streams);
    for (size_t i = 0; i < streams.size(); ++i)
      DM_HIP(hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking));
    DM_HIP(hipStreamCreateWithFlags(&origin, hipStreamNonBlocking));
  }
  events.resize(groups.size());
  for (auto& e : events) DM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  fork_ev.resize(1);
  DM_HIP(hipEventCreateWithFlags(&fork_ev[0], hipEventDisableTiming));
  if (opts.profile) {
    prof_start.resize(groups.size());
    prof_stop.resize(groups.size());
    for (auto& e : prof_start) DM_HIP(hipEventCreate(&e));
    for (auto& e : prof_stop) DM_HIP(hipEventCreate(&e));
    for (Group& G : groups)
      for (Group::Step& st : G.tsteps) {
        DM_HIP(hipEventCreate(&st.ev0));
        DM_HIP(hipEventCreate(&st.ev1));
      }
  }
  join_ev.resize(opts.num_streams);
  for (auto& e : join_ev) DM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return 0;
}

int Dag::build_device_tables() {
  // Host-side job tables for batched groups, then one upload.
  std::vector<unsigned char> blob;
  auto align = [&](size_t a) { while (blob.size() % a) blob.push_back(0); };
  for (Group& G : groups) {
    if (is_cplx(G.dtype)) continue;  // complex tasks need no job tables
    if (G.tasks.size() <= 1 && G.op != DM_OP_TRSM && G.op != DM_OP_GEMM) continue;
    if (G.op == DM_OP_GEMM) {
      const dm_task& t0 = tasks[G.tasks[0]].t;
      bool ta, tb;
      if (t0.op == DM_OP_GEMM) { ta = is_trans(t0.transa); tb = is_trans(t0.transb); }
      else { ta = is_trans(t0.transa); tb = !ta; }
      G.ta = ta;
      G.tb = tb;
      auto fill = [&](auto* dummy) -> int {
        typedef std::remove_pointer_t<decltype(dummy)> T;
        std::vector<GemmJob<T>> jobs;
        int maxM = 0, maxN = 0, maxK = 0;
        for (int ti : G.tasks) {
          const dm_task& t = tasks[ti].t;
          GemmJob<T> j{};
          if (t.op == DM_OP_GEMM) {
            j.A = (const T*)t.a; j.B = (const T*)t.b; j.C = (T*)t.c;
            j.M = t.m; j.N = t.n; j.K = t.k;
            j.lda = t.lda; j.ldb = t.ldb; j.ldc = t.ldc;
            j.tri = 0;
          } else {  // SYRK as masked GEMM of A with itself
            j.A = (const T*)t.a; j.B = (const T*)t.a; j.C = (T*)t.c;
            j.M = t.n; j.N = t.n; j.K = t.k;
            j.lda = t.lda; j.ldb = t.lda; j.ldc = t.ldc;
            j.tri = is_upper(t.uplo) ? 1 : 2;
          }
          j.alpha = (T)t.alpha;
          j.beta = (T)t.beta;
          if (j.K <= 0 || j.alpha == T(0)) { j.K = 0; j.alpha = T(0); }
          if (j.M <= 0 || j.N <= 0) continue;
          jobs.push_back(j);
          maxM = std::max(maxM, j.M); maxN = std::max(maxN, j.N); maxK = std::max(maxK, j.K);
        }
        G.njobs = (int)jobs.size();
        if (!G.njobs) return 0;
        G.vec = true;
        for (auto& j : jobs) G.vec = G.vec && gemm_job_vec_ok<T>(ta, tb, j);
        G.cfg = gemm_pick_cfg<T>(maxM, maxN, maxK, G.njobs, G.vec);
        if constexpr (std::is_same<T, float>::value) {
          if (opts.f32_update == 3 || opts.f32_update == 6) {
            bool split = true;
            for (auto& j : jobs) split = split && gemm_split_ok(ta, tb, j);
            if (split) G.cfg = gemm_split_cfg(opts.f32_update, maxM, maxN, G.njobs);
          }
        }
        G.total_wg = gemm_prepare_jobs<T>(jobs.data(), G.njobs, G.cfg);
        align(64);
        G.table_off = blob.size();
        const unsigned char* p = (const unsigned char*)jobs.data();
        blob.insert(blob.end(), p, p + jobs.size() * sizeof(GemmJob<T>));
        return 0;
      };
      int rc = G.dtype == DM_F64 ? fill((double*)nullptr) : fill((float*)nullptr);
      if (rc) return rc;
    } else if (G.op == DM_OP_TRSM) {
      const dm_task& t0 = tasks[G.tasks[0]].t;
      G.shape = trsm_shape(t0.side, t0.uplo, t0.transa, t0.diag);
      auto fill = [&](auto* dummy) -> int {
        typedef std::remove_pointer_t<decltype(dummy)> T;
        // one inverse-block buffer per distinct (A, L); job inv pointers
        // hold offsets into the inverse area until the allocation exists
        std::map<std::pair<const void*, int>, int> invidx;
        std::vector<TrsmJob<T>> jobs;
        for (int ti : G.tasks) {
          const dm_task& t = tasks[ti].t;
          const int L = G.shape.side_l ? t.m : t.n;
          const int NV = G.shape.side_l ? t.n : t.m;
          if (L <= 0 || NV <= 0) continue;
          auto key = std::make_pair(t.a, L);
          auto it = invidx.find(key);
          int ii;
          if (it == invidx.end()) {
            ii = (int)G.inv_src.size();
            invidx.emplace(key, ii);
            InvSrc src;
            src.A = t.a;
            src.lda = t.lda;
            src.L = L;
            src.off = inv_bytes;
            inv_bytes += sizeof(T) * (size_t)((L + TRSM_BS - 1) / TRSM_BS) * TRSM_BS * TRSM_BS;
            inv_bytes = (inv_bytes + 255) & ~(size_t)255;
            G.inv_src.push_back(src);
          } else {
            ii = it->second;
          }
          TrsmJob<T> j{};
          j.A = (const T*)t.a; j.B = (T*)t.c;
          j.inv = (const T*)(uintptr_t)G.inv_src[ii].off;  // patched after allocation
          j.L = L; j.NV = NV; j.lda = t.lda; j.ldb = t.ldc;
          j.alpha = (T)t.alpha;
          jobs.push_back(j);
        }
        G.batched_trsm = true;
        std::vector<TrsmStep<T>> steps;
        trsm_plan_steps<T>(G.shape, jobs, steps);
        for (TrsmStep<T>& st : steps) {
          Group::Step rec;
          rec.leaf = st.leaf;
          align(64);
          rec.table_off = blob.size();
          if (st.leaf) {
            rec.njobs = (int)st.tjobs.size();
            rec.maxL = st.maxL;
            for (const TrsmJob<T>& j : st.tjobs) rec.flops += (double)j.L * j.L * j.NV;
            rec.total_wg = st.total_wg;
            const unsigned char* p = (const unsigned char*)st.tjobs.data();
            blob.insert(blob.end(), p, p + st.tjobs.size() * sizeof(TrsmJob<T>));
          } else {
            int maxM = 0, maxN = 0, maxK = 0;
            for (const GemmJob<T>& g : st.gjobs) {
              maxM = std::max(maxM, g.M); maxN = std::max(maxN, g.N); maxK = std::max(maxK, g.K);
            }
            rec.njobs = (int)st.gjobs.size();
            for (const GemmJob<T>& g : st.gjobs) rec.flops += 2.0 * g.M * g.N * g.K;
            rec.ta = st.ta;
            rec.tb = st.tb;
            rec.vec = true;
            for (const GemmJob<T>& g : st.gjobs) rec.vec = rec.vec && gemm_job_vec_ok<T>(st.ta, st.tb, g);
            rec.cfg = gemm_pick_cfg<T>(maxM, maxN, maxK, rec.njobs, rec.vec);
            if constexpr (std::is_same<T, float>::value) {
              // the solve's updates take the same BF16-split arithmetic as
              // the factorization's when their operands allow it
              if ((opts.f32_update == 3 || opts.f32_update == 6) && maxN > 48) {
                bool split = true;
                for (const GemmJob<T>& g : st.gjobs) split = split && gemm_split_ok(st.ta, st.tb, g);
                if (split) rec.cfg = gemm_split_cfg(opts.f32_update, maxM, maxN, rec.njobs);
              }
            }
            rec.total_wg = gemm_prepare_jobs<T>(st.gjobs.data(), rec.njobs, rec.cfg);
            const unsigned char* p = (const unsigned char*)st.gjobs.data();
            blob.insert(blob.end(), p, p + st.gjobs.size() * sizeof(GemmJob<T>));
          }
          G.tsteps.push_back(rec);
        }
        return 0;
      };
      int rc = G.dtype == DM_F64 ? fill((double*)nullptr) : fill((float*)nullptr);
      if (rc) return rc;
    }
  }
  // per-task workspace for launches that solve through trsm_full
  // (POTRF recursion, unbatched TRSM/TRSV): allocated here so nothing is
  // allocated while a run is issued or captured.
  task_ws_off.assign(tasks.size(), (size_t)-1);
  size_t ws_bytes = 0;
  // LU panels share one workspace: they are ordered by their data (each
  // panel reads the updates of the previous one) and by the info word.
  size_t getrf_bytes = 0;
  for (const TaskRec& r : tasks)
    if (r.t.op == DM_OP_GETRF)
      getrf_bytes = std::max(getrf_bytes, r.t.dtype == DM_F64
                                              ? getrf_panel_ws_bytes<double>(r.t.m, r.t.n)
                                              : getrf_panel_ws_bytes<float>(r.t.m, r.t.n));
  if (getrf_bytes) {
    for (size_t i = 0; i < tasks.size(); ++i)
      if (tasks[i].t.op == DM_OP_GETRF) task_ws_off[i] = 256 + inv_bytes;
    ws_bytes = (getrf_bytes + 255) / 256 * 256;
  }
  for (Group& G : groups) {
    const bool needs = !is_cplx(G.dtype) &&
                       (G.op == DM_OP_TRSV || (G.op == DM_OP_TRSM && !G.batched_trsm));
    if (!needs) continue;
    const size_t per = G.dtype == DM_F64 ? trsm_ws_elems<double>() * sizeof(double)
                                         : trsm_ws_elems<float>() * sizeof(float);
    for (int ti : G.tasks) {
      task_ws_off[ti] = 256 + inv_bytes + ws_bytes;
      ws_bytes += per;
    }
  }
  // FP32 POTRF tasks factored in FP64 (opts.f64_diag): an n x n FP64 tile
  // each, so concurrent POTRFs of independent matrices never share one
  if (opts.f64_diag)
    for (size_t i = 0; i < tasks.size(); ++i) {
      const dm_task& t = tasks[i].t;
      if (t.op != DM_OP_POTRF || t.dtype != DM_F32) continue;
      task_ws_off[i] = 256 + inv_bytes + ws_bytes;
      ws_bytes += ((size_t)t.n * t.n * sizeof(double) + 255) / 256 * 256;
    }
  // device memory: [abort int (256 B)] [inverse blocks] [workspaces] [job tables]
  const size_t tab_base = 256 + inv_bytes + ws_bytes;
  const size_t total = tab_base + blob.size();
  DM_HIP(hipMalloc(&dev_mem, total));
  abort_dev = (int*)dev_mem;
  unsigned char* base = (unsigned char*)dev_mem;
  // patch inverse pointers inside the leaf tables of TRSM groups
  for (Group& G : groups) {
    if (G.op != DM_OP_TRSM) continue;
    for (Group::Step& st : G.tsteps) {
      if (!st.leaf) continue;
      if (G.dtype == DM_F64) {
        auto* jobs = (TrsmJob<double>*)(blob.data() + st.table_off);
        for (int i = 0; i < st.njobs; ++i)
          jobs[i].inv = (const double*)(base + 256 + (uintptr_t)jobs[i].inv);
      } else {
        auto* jobs = (TrsmJob<float>*)(blob.data() + st.table_off);
        for (int i = 0; i < st.njobs; ++i)
          jobs[i].inv = (const float*)(base + 256 + (uintptr_t)jobs[i].inv);
      }
    }
  }
  for (Group& G : groups) {
    G.table_dev = base + tab_base + G.table_off;
    for (InvSrc& s : G.inv_src) s.dev = base + 256 + s.off;
    for (Group::Step& st : G.tsteps) st.table_dev = base + tab_base + st.table_off;
  }
  if (!blob.empty()) DM_HIP(hipMemcpy(base + tab_base, blob.data(), blob.size(), hipMemcpyHostToDevice));
  DM_HIP(hipMemset(dev_mem, 0, 256));
  return 0;
}

// Complex tasks (each its own group): POTRF, TRSM, SYRK as herk!, GEMM with
// the real alpha/beta the tiled Cholesky and its solves use
// (src/array/cholesky.jl:23-26, src/array/trsm.jl:59-192).
template <typename R>
int Dag::launch_group_cplx(const Group& G, hipStream_t s) {
  for (int ti : G.tasks) {
    const dm_task& t = tasks[ti].t;
    const R al = (R)t.alpha, be = (R)t.beta;
    int rc = 0;
    switch (t.op) {
      case DM_OP_POTRF:
        rc = zpotrf_rec<R>(is_upper(t.uplo), t.n, (R*)t.c, t.ldc, t.info, 0, abort_dev, t.k, s);
        break;
      case DM_OP_TRSM:
        rc = ztrsm_call<R>(t.side, t.uplo, t.transa, t.diag, t.m, t.n, al, R(0), (const R*)t.a,
                           t.lda, (R*)t.c, t.ldc, abort_dev, s);
        break;
      case DM_OP_SYRK: {  // herk!(uplo, trans, alpha, A, beta, C)
        if ((t.alpha == 0.0 || t.k == 0) && t.beta == 1.0) break;  // BLAS quick return
        const bool tr = !is_notrans(t.transa);
        rc = zgemm_call<R>(tr ? 'C' : 'N', tr ? 'N' : 'C', t.n, t.n, t.k, al, R(0),
                           (const R*)t.a, t.lda, (const R*)t.a, t.lda, be, R(0), (R*)t.c, t.ldc,
                           is_upper(t.uplo) ? 1 : 2, 1, abort_dev, s);
        break;
      }
      case DM_OP_GEMM:
        rc = zgemm_call<R>(t.transa, t.transb, t.m, t.n, t.k, al, R(0), (const R*)t.a, t.lda,
                           (const R*)t.b, t.ldb, be, R(0), (R*)t.c, t.ldc, 0, 0, abort_dev, s);
        break;
      default:
        rc = arg_fail(0, "complex task op");
    }
    if (rc) return rc;
  }
  return 0;
}

template <typename T>
int Dag::launch_group_t(const Group& G, hipStream_t s) {
  const int* ab = abort_dev;
  // In a look-ahead schedule one side raises its waves' priority over the
  // other sharing the CUs: the critical stream's launches in plans with
  // factorization tasks (crit_prio), the bulk launches otherwise.
  const bool la = opts.lookahead > 0 && opts.num_streams >= 2;
  const int prio = la && (crit_prio ? G.crit == 0 : G.crit != 0) ? 1 : 0;
  if (G.op == DM_OP_GEMM) {
    if (G.njobs == 0) return 0;
    GemmJob<T> dummy{};
    dummy.prio = prio;
    return gemm_launch<T>(G.ta, G.tb, G.vec, G.cfg, (const GemmJob<T>*)G.table_dev, G.njobs, dummy,
                          G.total_wg, ab, s);
  }
  if (G.op == DM_OP_TRSM && G.batched_trsm) {
    for (const InvSrc& src : G.inv_src) {
      int rc = trinv_blocks_launch<T>(G.shape, (const T*)src.A, src.lda, src.L, (T*)src.dev, ab, s,
                                      prio);
      if (rc) return rc;
    }
    const bool prof = opts.profile && !capturing;
    for (const Group::Step& st : G.tsteps) {
      int rc;
      if (prof && st.ev0) DM_HIP(hipEventRecord(st.ev0, s));
      if (st.leaf) {
        TrsmJob<T> dummy{};
        dummy.L = st.maxL;
        dummy.prio = prio;
        rc = trsm_launch<T>(G.shape, (const TrsmJob<T>*)st.table_dev, st.njobs, dummy, st.total_wg,
                            ab, s);
      } else {
        GemmJob<T> dummy{};
        dummy.prio = prio;
        rc = gemm_launch<T>(st.ta, st.tb, st.vec, st.cfg, (const GemmJob<T>*)st.table_dev, st.njobs,
                            dummy, st.total_wg, ab, s);
      }
      if (rc) return rc;
      if (prof && st.ev1) DM_HIP(hipEventRecord(st.ev1, s));
    }
    return 0;
  }
  // Unbatched ops: one call per task in submission order.
  for (int ti : G.tasks) {
    const dm_task& t = tasks[ti].t;
    T* ws = task_ws_off[ti] == (size_t)-1 ? nullptr : (T*)((char*)dev_mem + task_ws_off[ti]);
    int rc = 0;
    switch (t.op) {
      case DM_OP_POTRF: {
        const int gofs = t.k;  // global row offset of the tile (set by the builders; 0 otherwise)
        if constexpr (std::is_same<T, float>::value) {
          if (opts.f64_diag) {  // ws: the task's FP64 tile
            rc = potrf_f64_of_f32(is_upper(t.uplo), t.n, (float*)t.c, t.ldc, (double*)ws, t.info,
                                  abort_dev, gofs, s);
            break;
          }
        }
        rc = potrf_full<T>(is_upper(t.uplo), t.n, (T*)t.c, t.ldc, t.info, abort_dev, gofs, ws, s);
        break;
      }
      case DM_OP_TRSM:
        rc = trsm_full<T>(t.side, t.uplo, t.transa, t.diag, t.m, t.n, (T)t.alpha, (const T*)t.a,
                          t.lda, (T*)t.c, t.ldc, ws, ab, s);
        break;
      case DM_OP_GEMM:
        rc = gemm_single<T>(is_trans(t.transa), is_trans(t.transb), t.m, t.n, t.k, (T)t.alpha,
                            (const T*)t.a, t.lda, (const T*)t.b, t.ldb, (T)t.beta, (T*)t.c, t.ldc,
                            0, ab, s);
        break;
      case DM_OP_SYRK: {
        const bool tr = is_trans(t.transa);
        rc = gemm_single<T>(tr, !tr, t.n, t.n, t.k, (T)t.alpha, (const T*)t.a, t.lda,
                            (const T*)t.a, t.lda, (T)t.beta, (T*)t.c, t.ldc,
                            is_upper(t.uplo) ? 1 : 2, ab, s);
        break;
      }
      case DM_OP_GEMV:
        rc = gemv_launch<T>(is_trans(t.transa), t.m, t.n, (T)t.alpha, (const T*)t.a, t.lda,
                            (const T*)t.b, t.ldb, (T)t.beta, (T*)t.c, t.ldc, ab, s);
        break;
      case DM_OP_GETRF:
        rc = getrf_panel<T>((T*)t.c, t.m, t.n, t.k, t.ldb, (int*)t.b, t.info, abort_dev,
                            t.transa == 'P' || t.transa == 'p', ws, s);
        break;
      case DM_OP_LASWP:
        rc = laswp_tiled<T>((T*)t.c, t.m, t.n, t.k, (const int*)t.b, t.lda, t.ldb, ab, s);
        break;
      case DM_OP_TRSV:
        if (t.ldc == 1)
          rc = trsm_full<T>('L', t.uplo, t.transa, t.diag, t.n, 1, T(1), (const T*)t.a, t.lda,
                            (T*)t.c, t.n, ws, ab, s);
        else
          rc = trsm_full<T>('R', t.uplo, is_trans(t.transa) ? 'N' : 'T', t.diag, 1, t.n, T(1),
                            (const T*)t.a, t.lda, (T*)t.c, t.ldc, ws, ab, s);
        break;
    }
    if (rc) return rc;
  }
  return 0;
}

int Dag::issue(hipStream_t o) {
  // reset the abort flag, fork, run groups, join
  int rc = fill_i32_launch(abort_dev, 0, 1, o);
  if (rc) return rc;
  if (pre_hook) {
    rc = pre_hook(this, o);
    if (rc) return rc;
  }
  // worker streams: the plan's own, else the caller's shared executor
  // streams; a one-stream eager plan runs on the caller's stream itself
  hipStream_t ss[8];
  const int ns = opts.num_streams;
  if (!streams.empty()) {
    for (int i = 0; i < ns; ++i) ss[i] = streams[i];
  } else if (ns == 1) {
    ss[0] = o;
  } else {
    rc = exec_streams(o, ns, ss);
    if (rc) return rc;
  }
  DM_HIP(hipEventRecord(fork_ev[0], o));
  for (int i = 0; i < ns; ++i)
    if (ss[i] != o) DM_HIP(hipStreamWaitEvent(ss[i], fork_ev[0], 0));
  for (int g : order) {
    const Group& G = groups[g];
    hipStream_t s = ss[G.stream];
    for (int w : G.waits) DM_HIP(hipStreamWaitEvent(s, events[w], 0));
    const bool prof = opts.profile && !capturing;
    if (prof) DM_HIP(hipEventRecord(prof_start[g], s));
    switch (G.dtype) {
      case DM_F64: rc = launch_group_t<double>(G, s); break;
      case DM_F32: rc = launch_group_t<float>(G, s); break;
      case DM_C64: rc = launch_group_cplx<double>(G, s); break;
      default: rc = launch_group_cplx<float>(G, s); break;
    }
    if (rc) return rc;
    if (prof) DM_HIP(hipEventRecord(prof_stop[g], s));
    DM_HIP(hipEventRecord(events[g], s));
  }
  for (int i = 0; i < ns; ++i) {
    if (ss[i] == o) continue;
    DM_HIP(hipEventRecord(join_ev[i], ss[i]));
    DM_HIP(hipStreamWaitEvent(o, join_ev[i], 0));
  }
  if (post_hook) {
    rc = post_hook(this, o);
    if (rc) return rc;
  }
  profiled = opts.profile && !capturing;
  return 0;
}

int Dag::capture() {
  // Capture, instantiate (and later launch) on the big-stack helper thread:
  // the runtime's graph walk recurses along dependency chains.
  return run_big_stack([this]() -> int {
    DM_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
    capturing = true;
    int rc = issue(origin);
    capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(origin, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    graph = g;
    DM_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    size_t nn = 0;
    (void)hipGraphGetNodes(graph, nullptr, &nn);
    graph_nodes = (int)nn;
    return 0;
  });
}

int Dag::run(hipStream_t s) {
  if (!planned) return arg_fail(1, "dag not planned");
  if (graph_exec) {
    hipGraphExec_t ex = graph_exec;
    return run_big_stack([ex, s]() -> int {
      DM_HIP(hipGraphLaunch(ex, s));
      return 0;
    });
  }
  // Eager issue: the origin stream is the caller's stream.
  return issue(s);
}

int Dag::launches() const {
  int n = 0;
  for (const Group& G : groups) {
    if (G.op == DM_OP_GEMM) n += G.njobs > 0 ? 1 : 0;
    else if (G.op == DM_OP_TRSM && G.batched_trsm) n += (int)(G.inv_src.size() + G.tsteps.size());
    else n += (int)G.tasks.size();
  }
  return n;
}

Dag::~Dag() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto& e : events) (void)hipEventDestroy(e);
  for (auto& e : fork_ev) (void)hipEventDestroy(e);
  for (auto& e : join_ev) (void)hipEventDestroy(e);
  for (auto& e : prof_start) (void)hipEventDestroy(e);
  for (auto& e : prof_stop) (void)hipEventDestroy(e);
  for (Group& G : groups)
    for (Group::Step& st : G.tsteps) {
      if (st.ev0) (void)hipEventDestroy(st.ev0);
      if (st.ev1) (void)hipEventDestroy(st.ev1);
    }
  // queued launches of this plan (on its own or on shared streams) may
  // still read its tables
  (void)hipDeviceSynchronize();
  for (auto& s : streams) (void)hipStreamDestroy(s);
  if (origin) (void)hipStreamDestroy(origin);
  if (dev_mem) (void)hipFree(dev_mem);
}

}  // namespace dm

using namespace dm;

extern "C" {

int dm_dag_create(dm_dag_t* dag) {
  if (!dag) return arg_fail(1, "dag is NULL");
  *dag = new dm_dag_s{std::unique_ptr<Dag>(new Dag())};
  return 0;
}

int dm_dag_destroy(dm_dag_t dag) {
  if (!dag) return 0;
  delete dag;
  return 0;
}

int dm_dag_add_task(dm_dag_t dag, const dm_task* t) {
  if (!dag) return arg_fail(1, "dag is NULL");
  if (!t) return arg_fail(2, "task is NULL");
  return dag->d->add(*t);
}

int dm_dag_edges(dm_dag_t dag, int* pairs, int* count) {
  if (!dag) return arg_fail(1, "dag is NULL");
  if (!count) return arg_fail(3, "count is NULL");
  const int cap = *count;
  int n = 0;
  for (size_t i = 0; i < dag->d->tasks.size(); ++i)
    for (int p : dag->d->tasks[i].preds) {
      if (pairs && n < cap) {
        pairs[2 * n] = p;
        pairs[2 * n + 1] = (int)i;
      }
      ++n;
    }
  *count = n;
  return 0;
}

int dm_dag_plan(dm_dag_t dag, const dm_plan_opts* opts) {
  if (!dag) return arg_fail(1, "dag is NULL");
  dm_plan_opts o{2, 1, 1, 1, 0};
  if (opts) o = *opts;
  int rc = ensure_init();
  if (rc) return rc;
  return dag->d->plan(o);
}

int dm_dag_run(dm_dag_t dag, dm_stream_t s) {
  if (!dag) return arg_fail(1, "dag is NULL");
  int rc = ensure_init();
  if (rc) return rc;
  return dag->d->run(stream_of(s));
}

int dm_dag_profile(dm_dag_t dag, dm_group_profile* out, int* count) {
  if (!dag) return arg_fail(1, "dag is NULL");
  if (!count) return arg_fail(3, "count is NULL");
  Dag& d = *dag->d;
  if (!d.profiled) return arg_fail(1, "no profiled run (plan with profile=1, no graph, then run)");
  int rc = ensure_init();
  if (rc) return rc;
  const int cap = *count;
  int n = 0;
  auto emit = [&](const Group& G, int op, int variant, double flops, hipEvent_t a,
                  hipEvent_t b) -> int {
    if (out && n < cap) {
      float ms = 0.f;
      DM_HIP(hipEventSynchronize(b));
      DM_HIP(hipEventElapsedTime(&ms, a, b));
      dm_group_profile& p = out[n];
      p.op = op;
      p.ntasks = (int)G.tasks.size();
      p.stream = G.stream;
      p.level = G.level;
      p.flops = flops;
      p.ms = ms;
      p.variant = variant;
    }
    ++n;
    return 0;
  };
  for (int g : d.order) {
    const Group& G = d.groups[g];
    int rc;
    if (G.op == DM_OP_TRSM && G.batched_trsm && !G.tsteps.empty() && G.tsteps[0].ev0) {
      // one record per launch of the batched solve: leaf solves (TRSM) and
      // the GEMM updates between them
      for (const Group::Step& st : G.tsteps) {
        const int variant = st.leaf ? -1 : st.cfg.id * 1000000 + st.cfg.bm * 1000 + st.cfg.bn;
        rc = emit(G, st.leaf ? DM_OP_TRSM : DM_OP_GEMM, variant, st.flops, st.ev0, st.ev1);
        if (rc) return rc;
      }
      continue;
    }
    const int variant = G.op == DM_OP_GEMM ? G.cfg.id * 1000000 + G.cfg.bm * 1000 + G.cfg.bn : -1;
    rc = emit(G, G.op, variant, G.flops, d.prof_start[g], d.prof_stop[g]);
    if (rc) return rc;
  }
  *count = n;
  return 0;
}

int dm_dag_get_stats(dm_dag_t dag, dm_dag_stats* st) {
  if (!dag) return arg_fail(1, "dag is NULL");
  if (!st) return arg_fail(2, "stats is NULL");
  const Dag& d = *dag->d;
  st->num_tasks = (int)d.tasks.size();
  int e = 0;
  for (auto& t : d.tasks) e += (int)t.preds.size();
  st->num_edges = e;
  st->num_groups = (int)d.groups.size();
  st->num_levels = d.num_levels;
  st->num_launches = d.planned ? d.launches() : 0;
  st->num_streams = d.opts.num_streams;
  st->graph = d.graph_exec ? 1 : 0;
  return 0;
}

}  // extern "C"
