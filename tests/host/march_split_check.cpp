// Host check of the march kernels' work split (formura_amd/csrc/fmr_stencil.hpp:
// march_split_of, march_item_range, march_runs -- the functions the kernels
// call).  Reads cases "ntile nx W T lx simds" from stdin; for each, prints
// "groups chunk max_iterations stored_once" where stored_once is 1 when the
// runs of all workgroups cover every (tile, plane) exactly once.
#include <cstdio>
#include <vector>

#include "fmr_stencil.hpp"

int main() {
  long long ntile;
  int nx, W, T, lx, simds;
  while (std::scanf("%lld %d %d %d %d %d", &ntile, &nx, &W, &T, &lx, &simds) == 6) {
    int groups = 0, chunk = 0;
    fmr::march_split_of(ntile, nx, W, T, lx, simds, groups, chunk);
    std::vector<int> seen((size_t)(ntile * nx), 0);
    long long max_iter = 0;
    bool ok = true;
    for (int g = 0; g < groups; ++g) {
      long long iters = 0;
      fmr::march_runs(g, groups, ntile, nx, chunk, [&](int tile, int x0, int xend) {
        if (tile < 0 || tile >= ntile || x0 < 0 || xend > nx || x0 >= xend) {
          ok = false;
          return;
        }
        iters += xend - x0 + 2 * T;
        for (int x = x0; x < xend; ++x) ++seen[(size_t)tile * nx + x];
      });
      if (iters > max_iter) max_iter = iters;
    }
    for (int c : seen) ok = ok && c == 1;
    std::printf("%d %d %lld %d\n", groups, chunk, max_iter, ok ? 1 : 0);
  }
  return 0;
}
