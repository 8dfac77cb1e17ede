// check_host.cpp -- drives the library's host-side validation and packing
// (boundaryvaluediffeq.jl_amd/csrc/bvp_host.cpp) with valid and invalid batches and
// options, built under AddressSanitizer / UndefinedBehaviorSanitizer by
// tests/test_host_sanitize.py.  Exit status 0 when every call returned the
// expected code and the packed arrays hold the expected values.
#include "../../boundaryvaluediffeq.jl_amd/csrc/bvp_host.h"

#include <cmath>
#include <cstdio>
#include <vector>

namespace {

int failures = 0;
void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAILED: %s\n", what);
        ++failures;
    }
}

// the kernels of the library: M = 1..4, 6 (wave) and 8, 16, 32 (block, LINEAR_TANH)
bool supports(int M, int fam) {
    if (fam == BVP_FAMILY_LINEAR_TANH) return (M >= 1 && M <= 4) || M == 6 || M == 8 || M == 16 || M == 32;
    if (M == 2) return fam == BVP_FAMILY_BRATU || fam == BVP_FAMILY_PENDULUM || fam == BVP_FAMILY_SINGPERT ||
                       fam == BVP_FAMILY_TURNING;
    if (M == 1) return fam == BVP_FAMILY_SCALAR_TU;
    if (M == 6) return fam == BVP_FAMILY_SWIRL;
    return false;
}

struct Inst {
    bvp_instance p{};
    std::vector<double> params, bct, bcv, u0, mesh, guess;
    std::vector<int32_t> bcc;
    void fix() {
        p.params = params.empty() ? nullptr : params.data();
        p.n_params = (int32_t)params.size();
        p.bc_t = bct.data();
        p.bc_comp = bcc.data();
        p.bc_val = bcv.data();
        p.n_bc = (int32_t)bct.size();
        p.u0 = u0.empty() ? nullptr : u0.data();
        p.mesh = mesh.empty() ? nullptr : mesh.data();
        p.guess = guess.empty() ? nullptr : guess.data();
    }
};

Inst bratu(double lam, int N) {
    Inst q;
    q.p.family = BVP_FAMILY_BRATU;
    q.p.M = 2;
    q.p.t0 = 0;
    q.p.t1 = 1;
    q.p.bc_kind = BVP_BC_TWO_POINT;
    q.p.N = N;
    q.params = {lam};
    q.bct = {0, 1};
    q.bcc = {0, 0};
    q.bcv = {0, 0};
    q.u0 = {0.1, -0.2};
    q.fix();
    return q;
}

int validate(std::vector<Inst> &v, int &M) {
    std::vector<bvp_instance> a;
    for (auto &q : v) {
        q.fix();
        a.push_back(q.p);
    }
    std::string err;
    return bvphost::validate_instances((int32_t)a.size(), a.data(), M, supports, err);
}

}  // namespace

int main() {
    int M = 0;
    // ---- valid batches and their packing
    {
        std::vector<Inst> v = {bratu(1.0, 7), bratu(2.0, 12), bratu(0.5, 12)};
        v[1].mesh.resize(13);
        for (int k = 0; k <= 12; ++k) v[1].mesh[k] = std::pow(k / 12.0, 2);
        v[1].guess.assign(26, 0.25);
        expect(validate(v, M) == BVP_OK && M == 2, "valid bratu batch");
        std::vector<bvp_instance> a;
        for (auto &q : v) a.push_back(q.p);
        bvphost::Packed pk;
        bvphost::pack_instances(3, a.data(), 2, pk);
        expect(pk.nodes == 8 + 13 + 13, "node count");
        expect(pk.off[0] == 0 && pk.off[1] == 8 && pk.off[2] == 21, "offsets");
        expect(pk.order[0] == 1 && pk.order[1] == 2 && pk.order[2] == 0, "longest first, stable");
        expect(pk.mesh[0] == 0.0 && pk.mesh[7] == 1.0 && pk.mesh[8 + 3] == 9.0 / 144.0, "meshes");
        expect(pk.y[0] == 0.1 && pk.y[1] == -0.2 && pk.y[2 * 8] == 0.25, "guesses");
        expect(pk.u0c[2] == 0.1 && pk.params.size() == 3 && pk.params[1] == 2.0 && pk.pcnt[2] == 1, "u0, params");
        std::string err;
        bvp_opts o{};
        o.method = BVP_MIRK4;
        o.abstol = 1e-6;
        o.max_num_subintervals = 3000;
        o.max_newton_iters = 1000;
        o.nlsolve_stages = 7;
        expect(bvphost::validate_opts(&o, 3, a.data(), 2, false, err) == BVP_OK, "default options");
        struct Bad { void (*set)(bvp_opts &); int code; const char *what; };
        const Bad bad[] = {
            {[](bvp_opts &x) { x.method = 5; }, BVP_EINVAL, "method"},
            {[](bvp_opts &x) { x.precision = 7; }, BVP_EINVAL, "precision"},
            {[](bvp_opts &x) { x.abstol = 0; }, BVP_EINVAL, "abstol 0"},
            {[](bvp_opts &x) { x.abstol = NAN; }, BVP_EINVAL, "abstol NaN"},
            {[](bvp_opts &x) { x.max_num_subintervals = 0; }, BVP_EINVAL, "max_num_subintervals"},
            {[](bvp_opts &x) { x.max_newton_iters = -1; }, BVP_EINVAL, "max_newton_iters"},
            {[](bvp_opts &x) { x.refine_iters = -1; }, BVP_EINVAL, "refine_iters"},
            {[](bvp_opts &x) { x.transfer_mode = 2; }, BVP_EINVAL, "transfer_mode"},
            {[](bvp_opts &x) { x.warm_start = 5; }, BVP_EINVAL, "warm_start"},
            {[](bvp_opts &x) { x.warm_start = BVP_WARM_MLP; }, BVP_EINVAL, "warm start without guesses"},
            {[](bvp_opts &x) { x.nlsolve_stages = 0; }, BVP_EINVAL, "stages 0"},
            {[](bvp_opts &x) { x.nlsolve_stages = 8; }, BVP_EINVAL, "stages 8"},
            {[](bvp_opts &x) { x.controller = 4; }, BVP_EINVAL, "controller"},
            {[](bvp_opts &x) { x.controller = BVP_CONTROL_GLOBAL; x.global_method = 2; }, BVP_EINVAL, "global_method"},
            {[](bvp_opts &x) { x.controller = BVP_CONTROL_GLOBAL; x.method = BVP_MIRK6; }, BVP_ENOTSUP, "HO with MIRK6"},
        };
        for (const Bad &b : bad) {
            bvp_opts x = o;
            b.set(x);
            expect(bvphost::validate_opts(&x, 3, a.data(), 2, false, err) == b.code, b.what);
        }
        bvp_opts x = o;
        x.nlsolve_stages = BVP_STAGE_TRUST_REGION;
        expect(bvphost::validate_opts(&x, 3, a.data(), 2, true, err) == BVP_ENOTSUP, "block kernels without NR stages");
        x = o;
        x.controller = BVP_CONTROL_GLOBAL;
        expect(bvphost::validate_opts(&x, 3, a.data(), 6, false, err) == BVP_ENOTSUP, "global control at M = 6");
        expect(bvphost::validate_opts(nullptr, 3, a.data(), 2, false, err) == BVP_EINVAL, "NULL opts");
    }
    {
        // a family without parameters, with params NULL; and the M = 6 family
        Inst s;
        s.p.family = BVP_FAMILY_SCALAR_TU;
        s.p.M = 1;
        s.p.t0 = 0;
        s.p.t1 = 1;
        s.p.bc_kind = BVP_BC_STANDARD;
        s.p.N = 5;
        s.bct = {0.5};
        s.bcc = {0};
        s.bcv = {1.0};
        s.u0 = {1.0};
        std::vector<Inst> v = {s};
        expect(validate(v, M) == BVP_OK && M == 1, "scalar family without parameters");
        std::vector<bvp_instance> a = {v[0].p};
        bvphost::Packed pk;
        bvphost::pack_instances(1, a.data(), 1, pk);
        expect(pk.params.empty() && pk.nodes == 6, "scalar packing");
        // a family that needs no parameters, given a positive count and no array: rejected
        // before the packing would read the array
        a[0].n_params = 2;
        a[0].params = nullptr;
        std::string err;
        expect(bvphost::validate_instances(1, a.data(), M, supports, err) == BVP_EINVAL, "params NULL, count 2");
    }
    // ---- invalid batches
    struct BadI { void (*set)(Inst &); int code; const char *what; };
    const BadI badi[] = {
        {[](Inst &q) { q.p.family = 99; }, BVP_ENOTSUP, "unknown family"},
        {[](Inst &q) { q.params.clear(); }, BVP_EINVAL, "missing parameters"},
        {[](Inst &q) { q.p.n_params = 1; q.params.clear(); q.fix(); q.p.n_params = 1; }, BVP_EINVAL, "params NULL"},
        {[](Inst &q) { q.bct.pop_back(); }, BVP_EINVAL, "BC rows"},
        {[](Inst &q) { q.p.bc_kind = 7; }, BVP_EINVAL, "bc_kind"},
        {[](Inst &q) { q.bcc[1] = 2; }, BVP_EINVAL, "bc_comp"},
        {[](Inst &q) { q.bct[1] = 0.5; }, BVP_EINVAL, "two-point row inside"},
        {[](Inst &q) { q.p.bc_kind = BVP_BC_STANDARD; q.bct[1] = 1.5; }, BVP_EINVAL, "BC point outside tspan"},
        {[](Inst &q) { q.p.N = 0; }, BVP_EINVAL, "N = 0"},
        {[](Inst &q) { q.p.t1 = 0; }, BVP_EINVAL, "t1 = t0"},
        {[](Inst &q) { q.p.t0 = NAN; q.bct[0] = NAN; }, BVP_EINVAL, "t0 NaN"},
        {[](Inst &q) { q.u0.clear(); }, BVP_EINVAL, "no guess, no u0"},
        {[](Inst &q) { q.mesh.assign(q.p.N + 1, 0.0); }, BVP_EINVAL, "mesh not increasing"},
        {[](Inst &q) { q.p.M = 3; }, BVP_EINVAL, "mixed M"},
    };
    for (const BadI &b : badi) {
        std::vector<Inst> v = {bratu(1.0, 6), bratu(1.0, 6)};
        b.set(v[1]);
        if (b.code == BVP_EINVAL && std::string(b.what) == "params NULL") {
            // fix() would re-point params; keep it NULL with a positive count
            v[0].fix();
            v[1].fix();
            std::vector<bvp_instance> a = {v[0].p, v[1].p};
            a[0].params = v[0].params.data();
            a[1].params = nullptr;
            a[1].n_params = 1;
            std::string err;
            expect(bvphost::validate_instances(2, a.data(), M, supports, err) == b.code, b.what);
            continue;
        }
        expect(validate(v, M) == b.code, b.what);
    }
    std::string err;
    expect(bvphost::validate_instances(0, nullptr, M, supports, err) == BVP_EINVAL, "empty batch");
    // ---- FIRK tableaus: order conditions B(p), collocation C(s) for Radau IIA / Lobatto
    // IIIA, and entries the reference tabulates (FIRK/radau_tableaus.jl:40-85,
    // FIRK/lobatto_tableaus.jl:146-166, 246-266, 293-323)
    {
        const int codes[] = {101, 102, 103, 105, 107, 202, 203, 204, 205, 302, 303, 304, 305, 402, 403, 404, 405};
        for (int code : codes) {
            bvphost::FirkTableau T;
            expect(bvphost::firk_tableau(code, T), "firk tableau exists");
            const int s = T.s;
            expect(s == code % 100, "stages");
            for (int k = 1; k <= T.order; ++k) {
                double v = 0;
                for (int j = 0; j < s; ++j) v += T.b[j] * std::pow(T.c[j], k - 1);
                expect(std::fabs(v - 1.0 / k) < 1e-13, "B(p)");
            }
            if (code / 100 <= 2)
                for (int k = 1; k <= s; ++k)
                    for (int i = 0; i < s; ++i) {
                        double v = 0;
                        for (int j = 0; j < s; ++j) v += T.A[i][j] * std::pow(T.c[j], k - 1);
                        expect(std::fabs(v - std::pow(T.c[i], k) / k) < 1e-13, "C(s)");
                    }
            // q' interpolates the stages: sum_k (k+1) Q[k][j] c_i^k = delta_ij
            for (int i = 0; i < s; ++i)
                for (int j = 0; j < s; ++j) {
                    double v = 0;
                    for (int k = 0; k < s; ++k) v += (k + 1) * T.Q[k][j] * std::pow(T.c[i], k);
                    expect(std::fabs(v - (i == j ? 1.0 : 0.0)) < 1e-11, "Q");
                }
        }
        bvphost::FirkTableau T;
        expect(!bvphost::firk_tableau(104, T) && !bvphost::firk_tableau(206, T) && !bvphost::firk_tableau(4, T),
               "unknown FIRK codes");
        const double r6 = std::sqrt(6.0), r21 = std::sqrt(21.0);
        bvphost::firk_tableau(103, T);
        expect(std::fabs(T.A[0][1] - (37.0 / 225 - 169 * r6 / 1800)) < 1e-14, "RadauIIa3 a12");
        expect(std::fabs(T.Q[1][1] - 3.320280554681775) < 1e-13, "RadauIIa3 q22");
        bvphost::firk_tableau(303, T);
        expect(std::fabs(T.A[2][1] - 5.0 / 6) < 1e-15 && std::fabs(T.A[0][2]) < 1e-15, "LobattoIIIb3");
        bvphost::firk_tableau(403, T);
        expect(std::fabs(T.A[1][2] + 1.0 / 12) < 1e-15 && T.tau == 0.7886751345948129, "LobattoIIIc3");
        bvphost::firk_tableau(405, T);
        expect(std::fabs(T.A[1][2] - (47 - 15 * r21) / 315) < 1e-14, "LobattoIIIc5 a23");
        expect(!T.adaptive == false, "LobattoIIIc5 adapts");
        bvphost::firk_tableau(402, T);
        expect(!T.adaptive, "LobattoIIIc2 does not adapt");
    }
    // ---- FIRK options
    {
        std::vector<Inst> v = {bratu(1.0, 7)};
        v[0].fix();
        std::vector<bvp_instance> a = {v[0].p};
        bvp_opts o{};
        o.abstol = 1e-6;
        o.max_num_subintervals = 3000;
        o.max_newton_iters = 1000;
        o.adaptive = 1;
        std::string err;
        o.method = 103;
        o.nlsolve_stages = 3;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_OK, "radau5 options");
        o.nlsolve_stages = 7;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_OK, "firk trust region");
        o.method = 103 + BVP_FIRK_NESTED;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_ENOTSUP, "nested trust region");
        o.method = 103;
        o.nlsolve_stages = 8;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_EINVAL, "firk stage bits");
        o.nlsolve_stages = 3;
        o.method = 101;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_ENOTSUP, "radau1 adaptive");
        o.adaptive = 0;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_OK, "radau1 fixed mesh");
        o.method = 106;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_EINVAL, "no radau6");
        o.method = 204;
        o.controller = BVP_CONTROL_GLOBAL;
        expect(bvphost::validate_opts(&o, 1, a.data(), 2, false, err) == BVP_ENOTSUP, "firk global error");
        o.controller = BVP_CONTROL_DEFECT;
        expect(bvphost::validate_opts(&o, 1, a.data(), 8, true, err) == BVP_ENOTSUP, "firk on the block kernels");
    }
    std::printf("%s: %d failures\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
