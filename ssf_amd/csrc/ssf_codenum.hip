// DOF numbering on the host: gtopology::gencodnum and give_code_numbers.
//
// Node flags per DOF (gnode::cn before numbering, GEFEL/gnode.cpp:55,128):
//   0   fixed (no unknown), stays 0
//   1   free: the next equation number
//   k>1 coupled group k: every DOF flagged k shares one equation number, taken
//       when the group is first met in the node ordering
//   <0  prescribed value index, kept as it is
// Nodes are visited in `ordering` (gtopology::ordering, the identity or a
// Cuthill-McKee order), DOFs of a node in their order
// (GEFEL/gtopology.cpp:1619-1665). Element-internal DOFs (cne == 1) do not occur
// for the hex8 / quad4 elements of this path.
#include <algorithm>
#include <vector>

#include "ssf_internal.hpp"

using namespace ssf;

extern "C" int ssf_gencodnum(int64_t nn, int64_t ndofn, int64_t *dof, const int64_t *ordering,
                             int64_t *ndof_out)
{
  SSF_CHECK(nn >= 0 && ndofn > 0 && (dof || nn == 0) && ndof_out, SSF_EINVAL,
            "ssf_gencodnum: bad argument");
  const int64_t total = nn * ndofn;
  // largest flag: groups 2..kmax need kmax - 1 slots (gtopology.cpp:1628-1641)
  int64_t kmax = 0;
  for (int64_t i = 0; i < total; i++) kmax = std::max(kmax, dof[i]);
  std::vector<int64_t> aux((size_t)std::max<int64_t>(kmax - 1, 0), -1);
  if (ordering) {
    std::vector<char> seen(nn, 0);
    for (int64_t i = 0; i < nn; i++) {
      SSF_CHECK(ordering[i] >= 0 && ordering[i] < nn && !seen[ordering[i]], SSF_EINVAL,
                "ssf_gencodnum: ordering is not a permutation of the nodes");
      seen[ordering[i]] = 1;
    }
  }
  int64_t next = 1;
  for (int64_t i = 0; i < nn; i++) {
    const int64_t ii = ordering ? ordering[i] : i;
    for (int64_t j = 0; j < ndofn; j++) {
      int64_t &k = dof[ii * ndofn + j];
      if (k <= 0) continue;  // fixed or prescribed
      if (k == 1) {
        k = next++;
      } else if (aux[k - 2] == -1) {
        aux[k - 2] = next;
        k = next++;
      } else {
        k = aux[k - 2];
      }
    }
  }
  *ndof_out = next - 1;
  return SSF_OK;
}

extern "C" int ssf_give_code_numbers(int64_t ne, int64_t nne, const int64_t *conn, int64_t nn,
                                     int64_t ndofn, const int64_t *dof, int64_t *cn)
{
  SSF_CHECK(ne >= 0 && nne > 0 && ndofn > 0 && (ne == 0 || (conn && dof && cn)), SSF_EINVAL,
            "ssf_give_code_numbers: bad argument");
  std::atomic<int> bad{0};
  parallel_for(ne, [&](int64_t b, int64_t e) {
    for (int64_t el = b; el < e; el++)
      for (int64_t a = 0; a < nne; a++) {
        const int64_t nd = conn[el * nne + a];
        if (nd < 0 || nd >= nn) {
          bad = 1;
          return;
        }
        for (int64_t j = 0; j < ndofn; j++) cn[(el * nne + a) * ndofn + j] = dof[nd * ndofn + j];
      }
  });
  SSF_CHECK(!bad.load(), SSF_EINVAL, "ssf_give_code_numbers: node index out of range");
  return SSF_OK;
}
