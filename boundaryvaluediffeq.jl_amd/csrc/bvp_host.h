// bvp_host.h -- host-side validation and packing of the C ABI's problem
// instances and options.  Plain C++ (no HIP): bvp_api.hip calls it, and it is
// built on its own under AddressSanitizer / UndefinedBehaviorSanitizer by the
// host checks.
#pragma once

#include <string>
#include <vector>

#include "../../include/bvp_mirk.h"

namespace bvphost {

// parameters a family needs (include/bvp_mirk.h, BVP_FAMILY_*)
int params_needed(int family, int M);

// range(t0; stop=t1, length=N+1) (Core/utils.jl:694-700), correctly rounded
// through an extended-precision intermediate
void uniform_mesh(double t0, double t1, int N, double *m);

// Tableau of a FIRK method code (BVP_RADAU_IIA(s), BVP_LOBATTO_IIIA/B/C(s)): A in Butcher
// order, b, c, Q (Q[k][j]: the coefficient of tau^(k+1) in int_0^tau l_j, l_j the
// Lagrange basis on the nodes), tau* (the defect sample point), order, and whether
// the method adapts the mesh.  Returns false for an unknown code.
struct FirkTableau {
    int s = 0, order = 0;
    bool adaptive = true;
    double A[7][7] = {}, b[7] = {}, c[7] = {}, Q[7][7] = {}, tau = 0;
};
bool firk_tableau(int code, FirkTableau &T);
inline bool is_firk(int method) { return method >= 100; }

// supports(M, family): a kernel of the library serves the pair
typedef bool (*SupportsFn)(int M, int family);

// Checks every instance of a batch (shared M, family known for M, parameters,
// boundary rows, tspan, mesh, guess).  Returns BVP_OK, BVP_EINVAL or BVP_ENOTSUP
// and the reason in err; M receives the batch's state dimension.
int validate_instances(int32_t n, const bvp_instance *inst, int &M, SupportsFn supports, std::string &err);

// Checks the options for a validated batch of state dimension M; blk: the batch
// runs on the block kernels.  Returns BVP_OK, BVP_EINVAL or BVP_ENOTSUP.
int validate_opts(const bvp_opts *o, int32_t n, const bvp_instance *inst, int M, bool blk, std::string &err);

// The packed problem description of a validated batch (structure of arrays per
// instance; initial meshes and guesses concatenated, instance i at node offset
// off[i]); order is the queue order, largest meshes first.
struct Packed {
    std::vector<int> fam, poff, pcnt, bck, N0, bcc, order;
    std::vector<double> t0, t1, params, bct, bcv, mesh, y, u0c;
    std::vector<long long> off;
    long long nodes = 0;
};
void pack_instances(int32_t n, const bvp_instance *inst, int M, Packed &p);

}  // namespace bvphost
