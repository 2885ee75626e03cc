// CPU harness: issue one tiled-Cholesky plan (dag.hip, tiled.hip) with HIP
// stubbed out (stubs.cpp), tracing every launch, event record and stream
// wait to $DM_HARNESS_TRACE.  The issue sequence is the one a hipGraph
// capture of the plan records (Dag::capture runs the same Dag::issue), so
// tests/test_capture_model.py can rebuild the captured dependency graph.
// Usage: capture_driver n nb num_streams lookahead U|L
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "dagger_mi355x.h"
int main(int argc, char** argv) {
  if (argc != 6) return 2;
  const long n = atol(argv[1]);
  const int nb = atoi(argv[2]), ns = atoi(argv[3]), la = atoi(argv[4]);
  const char uplo = argv[5][0];
  // a fake base: nothing dereferences tile memory in the stubbed launches
  dm_tiled t{(void*)(uintptr_t)(1ull << 40), n, n, nb, nb, DM_F64};
  std::vector<int> info(2);
  dm_dag_t d;
  if (dm_dag_build_cholesky(&d, uplo, &t, info.data())) return 3;
  dm_plan_opts o{ns, 1, la, 0, 0, 0};
  if (dm_dag_plan(d, &o)) return 4;
  if (dm_dag_run(d, nullptr)) return 5;
  dm_dag_stats st;
  dm_dag_get_stats(d, &st);
  printf("tasks %d groups %d levels %d launches %d\n", st.num_tasks, st.num_groups,
         st.num_levels, st.num_launches);
  dm_dag_destroy(d);
  return 0;
}
