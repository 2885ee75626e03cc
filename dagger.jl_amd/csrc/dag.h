// Datadeps executor internals (see dag.hip).
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <vector>
#include "dm_internal.h"

namespace dm {

struct Access {
  uintptr_t lo, hi;
  int mode;
};

struct TaskRec {
  dm_task t;
  std::vector<Access> acc;
  std::vector<int> preds, succs;
  double cost = 0;
  int level = 0, alap = 0, group = -1;
};

struct Region {
  uintptr_t lo = 0, hi = 0;
  int last_writer = -1;
  std::vector<int> readers;  // readers since last_writer
};

struct InvSrc {
  const void* A;
  int lda, L;
  size_t off;
  void* dev = nullptr;
};

struct Group {
  int level = 0, crit = 1, op = 0, dtype = DM_F64, stream = 0;
  std::vector<int> tasks;
  std::vector<int> waits;  // groups (on other streams) to wait for
  // batched GEMM/SYRK
  bool ta = false, tb = false, vec = false;
  GemmCfgId cfg{0, 128, 128};
  // batched TRSM: inverse blocks per distinct triangle, then the recursive
  // solve's launches (leaf solves and GEMM updates) in order
  TrsmShape shape{};
  bool batched_trsm = false;
  std::vector<InvSrc> inv_src;
  struct Step {
    bool leaf = true;
    int njobs = 0, maxL = 0;
    long total_wg = 0;
    bool ta = false, tb = false, vec = false;  // GEMM updates
    GemmCfgId cfg{0, 128, 128};
    size_t table_off = 0;
    void* table_dev = nullptr;
    double flops = 0;  // algorithmic flops of the step's launch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // profiled runs: around the launch
  };
  std::vector<Step> tsteps;
  // common
  double flops = 0;  // algorithmic flops of the group's tasks
  int njobs = 0;
  long total_wg = 0;
  size_t table_off = 0;
  void* table_dev = nullptr;
};

// Plans above this many flops issue eagerly rather than through a graph.
constexpr double kGraphMaxFlops = 2e11;

struct Dag {
  std::vector<TaskRec> tasks;
  std::vector<Region> regions;
  std::multimap<uintptr_t, int> regions_by_lo;
  uintptr_t max_len = 0;

  bool planned = false;
  dm_plan_opts opts{1, 0, 0, 0};
  // which side of a look-ahead schedule raises its waves' priority: the
  // critical stream's launches (plans with factorization tasks; set by
  // plan()) or the bulk launches
  bool crit_prio = true;
  int num_levels = 0;
  std::vector<Group> groups;
  std::vector<int> order;

  // device state owned by the plan
  void* dev_mem = nullptr;
  int* abort_dev = nullptr;  // 0, or the global order of the first failed pivot
  size_t inv_bytes = 0;
  std::vector<size_t> task_ws_off;  // byte offset in dev_mem, or SIZE_MAX
  std::vector<hipStream_t> streams;
  hipStream_t origin = nullptr;
  std::vector<hipEvent_t> events, fork_ev, join_ev;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_nodes = 0;
  bool capturing = false;
  // profiling (opts.profile, eager runs): events around each group
  std::vector<hipEvent_t> prof_start, prof_stop;
  bool profiled = false;
  // Extra work issued on the origin stream before the groups fork onto
  // their streams (pre_hook) and after they join (post_hook), in every run
  // (captured too).
  std::function<int(Dag*, hipStream_t)> pre_hook, post_hook;

  int add(const dm_task& t);
  int plan(const dm_plan_opts& o);
  int run(hipStream_t s);
  int launches() const;
  ~Dag();

 private:
  int build_device_tables();
  int create_streams(bool eager);
  int capture();
  int issue(hipStream_t origin_stream);
  template <typename T>
  int launch_group_t(const Group& G, hipStream_t s);
  template <typename R>
  int launch_group_cplx(const Group& G, hipStream_t s);
};

}  // namespace dm

struct dm_dag_s {
  std::unique_ptr<dm::Dag> d;
};
