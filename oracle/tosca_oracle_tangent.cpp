// tosca_oracle_tangent.cpp -- exact FP64 tangent (JVP) of the oracle's right-hand side.
//
// TEST INFRASTRUCTURE ONLY (see tosca_oracle.cpp).  The FormU / UpdateNut / Buoyancy part of
// tosca_oracle.cpp is written over a scalar type Real; this file compiles that same code with Real
// = Dual, a forward-mode dual number (value, tangent), so every product, quotient, square root and
// power of the oracle carries its exact derivative along.  The value channel performs the oracle's
// double operations in the oracle's order, so it reproduces the double oracle bit for bit.
//
// Two functions are differentiated.  The right-hand side the adjoint of the library is defined on:
//   R(ucont, ucat, theta, Cs) = FormU(ucont, ucat, nu_t, scale, formMode 0) + Buoyancy(theta, b)
//   nu_t = UpdateNut(ucat, Cs)   (interior Cs Delta^2 |S| plus the UpdateNutBCs ghosts)
// with every input a full DMDA local array (ghost cells included) and R on the owned faces; and the
// explicit Wray Runge-Kutta predictor over several stages (oracle_rk3w_jvp below), the function of
// the library's checkpointed predictor adjoint.

#include <cmath>
#include <vector>

struct Dual {
    double v, d;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator+(Dual a, double b) { return a + Dual(b); }
inline Dual operator+(double a, Dual b) { return Dual(a) + b; }
inline Dual operator-(Dual a, double b) { return a - Dual(b); }
inline Dual operator-(double a, Dual b) { return Dual(a) - b; }
inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
inline Dual operator/(double a, Dual b) { return Dual(a) / b; }
inline Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
inline Dual &operator-=(Dual &a, Dual b) { a = a - b; return a; }
inline Dual &operator*=(Dual &a, Dual b) { a = a * b; return a; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<(Dual a, double b) { return a.v < b; }
inline bool operator>(Dual a, double b) { return a.v > b; }
inline bool operator<(double a, Dual b) { return a < b.v; }
inline bool operator>(double a, Dual b) { return a > b.v; }
inline bool operator==(Dual a, double b) { return a.v == b; }
inline bool operator!=(Dual a, double b) { return a.v != b; }
// d sqrt(x) = dx / (2 sqrt(x)); at x = 0 (|S| of a uniform flow) the one-sided derivative is
// infinite and the tangent is taken as 0, the subgradient the adjoint uses as well
inline Dual sqrt(Dual a) {
    const double r = std::sqrt(a.v);
    return {r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
// d x^p = p x^(p-1) dx
inline Dual pow(Dual a, double p) {
    const double r = std::pow(a.v, p);
    return {r, a.d == 0.0 ? 0.0 : p * std::pow(a.v, p - 1.0) * a.d};
}

using Real = Dual;
#define ORACLE_TANGENT 1
struct MeshD {
    const Dual *csi, *eta, *zet, *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
    const Dual *aj, *iaj, *jaj, *kaj, *nvert, *meshtag;
};
#define ORACLE_MESH MeshD
#define ORACLE_API(name) name##_dual
#include "tosca_oracle.cpp"

namespace {
std::vector<Dual> lift(const double *v, const double *d, size_t n) {
    std::vector<Dual> r(n);
    for (size_t q = 0; q < n; q++) r[q] = Dual(v ? v[q] : 0.0, d ? d[q] : 0.0);
    return r;
}
}  // namespace

extern "C" {

// R and its tangent dR = J (d_ucont, d_ucat, d_theta, d_cs) for the function described at the top.
// Local arrays: ucont/ucat (3 per point), theta/cs (1 per point); R and dR: owned-range AoS.
// A null tangent pointer is a zero direction.
int oracle_rhs_jvp(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                   const double *ucat, const double *theta, const double *cs, const double *d_ucont,
                   const double *d_ucat, const double *d_theta, const double *d_cs, double scale, double buoy_scale,
                   double *R, double *dR) {
    if (!info || !m || !p) return TOSCA_ERR_ARG;
    const size_t n = (size_t)info->gxm * info->gym * info->gzm, ng = (size_t)info->mx * info->my * info->mz;
    const double *mv[18] = {m->csi, m->eta, m->zet, m->icsi, m->ieta, m->izet, m->jcsi, m->jeta, m->jzet,
                            m->kcsi, m->keta, m->kzet, m->aj, m->iaj, m->jaj, m->kaj, m->nvert, m->meshtag};
    std::vector<Dual> ms[18];
    for (int q = 0; q < 18; q++) ms[q] = lift(mv[q], nullptr, (q < 12 ? 3 : 1) * n);
    MeshD md{ms[0].data(), ms[1].data(), ms[2].data(), ms[3].data(), ms[4].data(), ms[5].data(),
             ms[6].data(), ms[7].data(), ms[8].data(), ms[9].data(), ms[10].data(), ms[11].data(),
             ms[12].data(), ms[13].data(), ms[14].data(), ms[15].data(), ms[16].data(), ms[17].data()};
    std::vector<Dual> V = lift(ucont, d_ucont, 3 * n), U = lift(ucat, d_ucat, 3 * n), T = lift(theta, d_theta, n),
                      C = lift(cs, d_cs, n);
    std::vector<Dual> nut(n), rhs(3 * ng, Dual(0.0)), bt(3 * ng, Dual(0.0));
    int rc = oracle_update_nut_dual(info, &md, p, U.data(), C.data(), nut.data(), nullptr);
    if (rc) return rc;
    rc = oracle_formu_dual(info, &md, p, V.data(), U.data(), nut.data(), rhs.data(), scale, 0);
    if (rc) return rc;
    rc = oracle_buoyancy_dual(info, &md, p, T.data(), buoy_scale, bt.data(), nullptr);
    if (rc) return rc;
    for (size_t q = 0; q < 3 * ng; q++) {
        const Dual r = rhs[q] + bt[q];
        if (R) R[q] = r.v;
        if (dR) dR[q] = r.d;
    }
    return TOSCA_OK;
}

// The explicit Wray Runge-Kutta predictor over n_stages stages (oracle_ueqn_rk3w, the sources,
// ucat's non-periodic boundary cells and nu_t held fixed) as a function of the initial Ucont (the
// owned range of the local ucont, a global vector) and of the local nu_t array, and its tangent:
// ucont_f / d_ucont_f (global) = the final Ucont and J (d_ucont, d_nut).  d_ucont: global
// direction; d_nut: local direction (null: zero).  ucont / ucat are not modified.
int oracle_rk3w_jvp(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                    const double *ucat, const double *nut, const double *dp, const double *btheta,
                    const double *btheta_o, int ab2, const double *farm, double dt, int n_stages,
                    const double *d_ucont, const double *d_nut, int form_mode, double *ucont_f, double *d_ucont_f) {
    if (!info || !m || !p || !ucont || !ucat || !dp) return TOSCA_ERR_ARG;
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    const size_t n = (size_t)info->gxm * info->gym * info->gzm, ng = (size_t)info->mx * info->my * info->mz;
    const double *mv[18] = {m->csi, m->eta, m->zet, m->icsi, m->ieta, m->izet, m->jcsi, m->jeta, m->jzet,
                            m->kcsi, m->keta, m->kzet, m->aj, m->iaj, m->jaj, m->kaj, m->nvert, m->meshtag};
    std::vector<Dual> ms[18];
    for (int q = 0; q < 18; q++) ms[q] = lift(mv[q], nullptr, (q < 12 ? 3 : 1) * n);
    MeshD md{ms[0].data(), ms[1].data(), ms[2].data(), ms[3].data(), ms[4].data(), ms[5].data(),
             ms[6].data(), ms[7].data(), ms[8].data(), ms[9].data(), ms[10].data(), ms[11].data(),
             ms[12].data(), ms[13].data(), ms[14].data(), ms[15].data(), ms[16].data(), ms[17].data()};
    std::vector<Dual> V = lift(ucont, nullptr, 3 * n), U = lift(ucat, nullptr, 3 * n);
    if (d_ucont) {                 // the direction on the owned range (the initial global Ucont)
        for (int64_t k = 0; k < l.mz; k++)
            for (int64_t j = 0; j < l.my; j++)
                for (int64_t i = 0; i < l.mx; i++)
                    for (int c = 0; c < 3; c++) V[3 * l.L(k, j, i) + c].d = d_ucont[3 * l.G(k, j, i) + c];
    }
    std::vector<Dual> NU = lift(nut, d_nut, n);
    std::vector<Dual> DP = lift(dp, nullptr, 3 * ng), BT = lift(btheta, nullptr, 3 * ng),
                      BTO = lift(btheta_o, nullptr, 3 * ng), FA = lift(farm, nullptr, 3 * ng);
    std::vector<Dual> uo(3 * ng), rhs(3 * ng), rhs_o(3 * ng);
    rc = oracle_ueqn_rk3w_dual(info, &md, p, V.data(), U.data(), nut ? NU.data() : nullptr, DP.data(),
                               btheta ? BT.data() : nullptr, btheta_o ? BTO.data() : nullptr, ab2,
                               farm ? FA.data() : nullptr, dt, n_stages, uo.data(), rhs.data(), rhs_o.data(),
                               form_mode);
    if (rc) return rc;
    for (int64_t k = 0; k < l.mz; k++)
        for (int64_t j = 0; j < l.my; j++)
            for (int64_t i = 0; i < l.mx; i++)
                for (int c = 0; c < 3; c++) {
                    const Dual v = V[3 * l.L(k, j, i) + c];
                    if (ucont_f) ucont_f[3 * l.G(k, j, i) + c] = v.v;
                    if (d_ucont_f) d_ucont_f[3 * l.G(k, j, i) + c] = v.d;
                }
    return TOSCA_OK;
}

}  // extern "C"
