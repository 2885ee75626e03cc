// CPU harness for the native planner: builds the tiled Cholesky / potrs /
// trsm / gemm DAGs under every executor option set through the real C ABI
// (dag.hip, tiled.hip) with HIP and the device launches stubbed out
// (stubs.cpp), and prints the plan statistics.  Run under ASan/UBSan by
// tests/test_planner_host.py.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "dagger_mi355x.h"
int main() {
  std::vector<double> slab(1536 * 1536), bslab(1536 * 64);
  int info[2];
  dm_tiled t{slab.data(), 1536, 1536, 256, 256, DM_F64};
  dm_tiled b{bslab.data(), 1536, 64, 256, 256, DM_F64};
  dm_tiled rag{slab.data(), 1000, 1000, 128, 128, DM_F64};
  dm_plan_opts cfgs[] = {{1, 0, 0, 0, 0, 0}, {2, 1, 1, 0, 0, 0}, {4, 1, 0, 0, 0, 0},
                         {3, 1, 2, 1, 0, 0}, {2, 1, 2, 1, 0, 0}, {2, 1, 1, 0, 1, 0}};
  for (auto& o : cfgs) {
    for (int uplo = 0; uplo < 2; ++uplo) {
      const char u = uplo ? 'L' : 'U';
      if (dm_cholesky_tiled(u, &t, info, info + 1, &o, nullptr)) return 1;
      if (dm_cholesky_tiled(u, &rag, info, info + 1, &o, nullptr)) return 2;
      if (dm_potrs_tiled(u, &t, &b, &o, nullptr)) return 3;
      dm_dag_t d;
      if (dm_dag_build_cholesky(&d, u, &t, info)) return 4;
      if (dm_dag_plan(d, &o)) return 5;
      if (dm_dag_run(d, nullptr)) return 6;
      dm_dag_stats st;
      dm_dag_get_stats(d, &st);
      // graphs are captured only for plans of at most two streams
      if (st.graph != (o.use_graph && o.num_streams <= 2 ? 1 : 0)) return 10;
      printf("opts %d%d%d%d%d %c: tasks %d edges %d groups %d levels %d launches %d\n",
             o.num_streams, o.batch, o.lookahead, o.use_graph, o.profile, u, st.num_tasks,
             st.num_edges, st.num_groups, st.num_levels, st.num_launches);
      if (o.profile) {
        dm_group_profile p[256];
        int cnt = 256;
        if (dm_dag_profile(d, p, &cnt)) return 7;
      }
      dm_dag_destroy(d);
    }
    if (dm_gemm_tiled('N', 'N', -1.0, &t, &b, 1.0, &b, &o, nullptr)) return 8;
    // LU with and without pivoting, square and ragged
    std::vector<int> ipiv(1536);
    for (int piv = 0; piv < 2; ++piv) {
      if (dm_getrf_tiled(piv, &t, ipiv.data(), info, &o, nullptr)) return 12;
      if (dm_getrf_tiled(piv, &rag, ipiv.data(), info + 1, &o, nullptr)) return 13;
    }
    if (dm_tiled_laswp(&b, ipiv.data(), 1536, nullptr)) return 14;
    for (char side : {'L', 'R'})
      for (char up : {'U', 'L'})
        for (char tr : {'N', 'T'}) {
          dm_tiled bb = side == 'L' ? b : dm_tiled{bslab.data(), 64, 1536, 256, 256, DM_F64};
          if (dm_trsm_tiled(side, up, tr, 'N', 1.0, &t, &bb, &o, nullptr)) return 9;
        }
  }
  // FP32 tiles with the BF16-split update (upper: k-contiguous operands)
  std::vector<float> fslab(1536 * 1536);
  dm_tiled tf{fslab.data(), 1536, 1536, 256, 256, DM_F32};
  for (int upd : {0, 3, 6}) {
    dm_plan_opts o{2, 1, 1, 0, 0, upd};
    for (char u : {'U', 'L'})
      if (dm_cholesky_tiled(u, &tf, info, info + 1, &o, nullptr)) return 11;
  }
  dm_plan_cache_clear();
  printf("ok\n");
  return 0;
}
