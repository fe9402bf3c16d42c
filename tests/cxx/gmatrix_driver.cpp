// A host C++ caller of the boundary: the gmatrix_gpu mirror (ssf_amd/host) driven as
// SIFEL's element loop drives gmatrix: initiate from code numbers, localized per
// element, solve_system, gmxv. Test infrastructure.
//   gmatrix_driver --layout          the C ABI struct layouts as JSON
//   gmatrix_driver IN OUT            IN: int64 ts, n, ne, nd, ni; double res;
//                                    int64 cn[ne nd]; double lm[ne nd nd]; double rhs[n]
//                                    OUT: int64 status, ani, negm; double ares;
//                                    double x[n], Ax[n], a[negm]
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gmatrix_gpu.hpp"

template <typename T>
static bool rd(FILE *f, T *p, size_t n)
{
  return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char **argv)
{
  if (argc == 2 && std::strcmp(argv[1], "--layout") == 0) {
    std::printf("{\"ssf_solve_params\": {\"size\": %zu, \"linsolver\": %zu, \"prec\": %zu, "
                "\"ni\": %zu, \"res\": %zu, \"zero\": %zu, \"iv\": %zu, \"precision\": %zu, "
                "\"inner_res\": %zu, \"inner_ni\": %zu, \"max_outer\": %zu}, ",
                sizeof(ssf_solve_params), offsetof(ssf_solve_params, linsolver),
                offsetof(ssf_solve_params, prec), offsetof(ssf_solve_params, ni),
                offsetof(ssf_solve_params, res), offsetof(ssf_solve_params, zero),
                offsetof(ssf_solve_params, iv), offsetof(ssf_solve_params, precision),
                offsetof(ssf_solve_params, inner_res), offsetof(ssf_solve_params, inner_ni),
                offsetof(ssf_solve_params, max_outer));
    std::printf("\"ssf_solve_info\": {\"size\": %zu, \"ani\": %zu, \"ares\": %zu, \"status\": %zu, "
                "\"pad\": %zu, \"iterations\": %zu, \"outer\": %zu, \"rel_residual\": %zu}}\n",
                sizeof(ssf_solve_info), offsetof(ssf_solve_info, ani),
                offsetof(ssf_solve_info, ares), offsetof(ssf_solve_info, status),
                offsetof(ssf_solve_info, pad), offsetof(ssf_solve_info, iterations),
                offsetof(ssf_solve_info, outer), offsetof(ssf_solve_info, rel_residual));
    return 0;
  }
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s --layout | IN OUT\n", argv[0]);
    return 2;
  }
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  long h[5];
  double res;
  if (!rd(f, h, 5) || !rd(f, &res, 1)) return 2;
  const long ts = h[0], n = h[1], ne = h[2], nd = h[3], ni = h[4];
  std::vector<long> cn(ne * nd);
  std::vector<double> lm(ne * nd * nd), rhs(n);
  if (!rd(f, cn.data(), cn.size()) || !rd(f, lm.data(), lm.size()) || !rd(f, rhs.data(), n))
    return 2;
  std::fclose(f);

  ssf_host::gmatrix_gpu G(0);
  G.ni = ni;
  G.res = res;
  G.prec = SSF_PREC_DIAG;
  G.initiate((int)ts, n, ne, nd, cn.data());
  for (long e = 0; e < ne; e++) G.localized(&lm[e * nd * nd], &cn[e * nd], nd, e);
  std::vector<double> x(n, 0.0), ax(n);
  G.solve_system(x.data(), rhs.data());
  G.gmxv(x.data(), ax.data());
  std::vector<double> a(G.negm());
  G.give_values(a.data());
  if (!G.ok()) return 1;
  FILE *o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  const long out_i[3] = {G.status, G.ani, G.negm()};
  std::fwrite(out_i, sizeof(long), 3, o);
  std::fwrite(&G.ares, sizeof(double), 1, o);
  std::fwrite(x.data(), sizeof(double), n, o);
  std::fwrite(ax.data(), sizeof(double), n, o);
  std::fwrite(a.data(), sizeof(double), a.size(), o);
  std::fclose(o);
  return 0;
}
