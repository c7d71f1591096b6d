// Checks of the plug points (include/euler/plugins.hpp), run by
// tests/test_gpu_cpp_api.py:
//   * the generic kernels with the reference's own solver, EOS and mesh (through
//     derived types, so that the generic overloads are chosen) equal the library's
//     tiled kernels bitwise: godunov, cons_to_prim, prim_to_cons, time_step;
//   * a caller-defined Riemann solver (Rusanov), EOS (stiffened gas) and mesh type
//     run on the GPU and equal a host evaluation of the reference's loops through
//     the same functors, bitwise, in FP64 and FP32.
// One line per check; exit status 0 iff all pass.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include <euler/plugins.hpp>

namespace {

int g_failures = 0;

void expect(bool ok, std::string const& what)
{
    std::cout << (ok ? "PASS " : "FAIL ") << what << '\n' << std::flush;
    if (!ok) {
        ++g_failures;
    }
}

// The reference's types under other names: same behaviour, generic overloads.
struct hllc_again : eb::hllc
{
};

template <class T>
struct perfect_gas_again : eb::PerfectGas<T>
{
    EB_FUNCTION explicit perfect_gas_again(T g) : eb::PerfectGas<T>(g) {}
};

template <class T>
struct mesh_again : eb::UniformMesh3d<T>
{
    EB_FUNCTION mesh_again(T a, T b, T c) : eb::UniformMesh3d<T>(a, b, c) {}
};

// Rusanov (local Lax-Friedrichs) flux: 0.5 (F(qL) + F(qR)) - 0.5 s (U(qR) - U(qL)),
// s = max(|u_n| + c) over both sides.
struct rusanov
{
    template <std::size_t D, class T, class EOS>
    EB_FUNCTION eb::EulerFlux<T> operator()(std::integral_constant<std::size_t, D> dir, EOS const& eos,
                                            eb::EulerPrim<T> const& qL, eb::EulerPrim<T> const& qR) const noexcept
    {
        EB_IEEE_ORDER
        T const unL = eb::get(dir, qL);
        T const unR = eb::get(dir, qR);
        T const aL = fabs(unL) + eos.speed_of_sound(qL.d, qL.p);
        T const aR = fabs(unR) + eos.speed_of_sound(qR.d, qR.p);
        T const s = aL < aR ? aR : aL;
        eb::EulerCons<T> const uL = eb::to_cons(qL, eos.internal_energy(qL.d, qL.p));
        eb::EulerCons<T> const uR = eb::to_cons(qR, eos.internal_energy(qR.d, qR.p));
        auto phys = [&](eb::EulerPrim<T> const& q, eb::EulerCons<T> const& u, T un) {
            eb::EulerFlux<T> f {u.d * un, (u.e + q.p) * un, u.mx0 * un, u.mx1 * un, u.mx2 * un};
            if constexpr (D == 0) {
                f.mx0 += q.p;
            } else if constexpr (D == 1) {
                f.mx1 += q.p;
            } else {
                f.mx2 += q.p;
            }
            return f;
        };
        eb::EulerFlux<T> const fL = phys(qL, uL, unL);
        eb::EulerFlux<T> const fR = phys(qR, uR, unR);
        T const h = T(0.5);
        return {h * (fL.d + fR.d) - h * s * (uR.d - uL.d), h * (fL.e + fR.e) - h * s * (uR.e - uL.e),
                h * (fL.mx0 + fR.mx0) - h * s * (uR.mx0 - uL.mx0), h * (fL.mx1 + fR.mx1) - h * s * (uR.mx1 - uL.mx1),
                h * (fL.mx2 + fR.mx2) - h * s * (uR.mx2 - uL.mx2)};
    }
};

// Stiffened gas: p + p_inf behaves as a perfect gas.
template <class T>
struct stiffened_gas
{
    T gamma;
    T p_inf;

    EB_FUNCTION T speed_of_sound(T d, T p) const noexcept
    {
        EB_IEEE_ORDER
        return sqrt(gamma * (p + p_inf) / d);
    }
    EB_FUNCTION T internal_energy(T /*d*/, T p) const noexcept
    {
        EB_IEEE_ORDER
        return (p + gamma * p_inf) / (gamma - 1);
    }
    EB_FUNCTION T pressure(T /*d*/, T eint) const noexcept
    {
        EB_IEEE_ORDER
        return (gamma - 1) * eint - gamma * p_inf;
    }
};

// A caller's own mesh type (cell widths, face areas, volume).
template <class T>
struct box_mesh
{
    T h0, h1, h2;
    EB_FUNCTION T dx0() const noexcept { return h0; }
    EB_FUNCTION T dx1() const noexcept { return h1; }
    EB_FUNCTION T dx2() const noexcept { return h2; }
    EB_FUNCTION T ds0() const noexcept { EB_IEEE_ORDER return h1 * h2; }
    EB_FUNCTION T ds1() const noexcept { EB_IEEE_ORDER return h2 * h0; }
    EB_FUNCTION T ds2() const noexcept { EB_IEEE_ORDER return h0 * h1; }
    EB_FUNCTION T dv() const noexcept { EB_IEEE_ORDER return h0 * h1 * h2; }
};

template <class T>
using view_t = eb::mdspan<T, eb::dextents<int, 3>, eb::layout_left>;

template <class T>
struct State
{
    eb::EulerPrimArrays<eb::View<T>> pa;
    eb::EulerConsArrays<eb::View<T>> ca;
    eb::EulerPrimArrays<view_t<T>> prim;
    eb::EulerConsArrays<view_t<T>> cons;

    State(eb::HipExecSpace const& ex, int n0, int n1, int n2)
        : pa(eb::create_prim_arrays_1d<T>(ex, std::size_t(n0) * n1 * n2)),
          ca(eb::create_cons_arrays_1d<T>(ex, std::size_t(n0) * n1 * n2)),
          prim(eb::to_mdspan<view_t<T>>(pa, n0, n1, n2)), cons(eb::to_mdspan<view_t<T>>(ca, n0, n1, n2))
    {
    }
};

template <class T>
std::vector<std::vector<T>> host_prim(eb::HipExecSpace const& ex, eb::EulerPrimArrays<view_t<T>> const& a)
{
    return {eb::to_host(ex, a.d), eb::to_host(ex, a.p), eb::to_host(ex, a.ux0), eb::to_host(ex, a.ux1),
            eb::to_host(ex, a.ux2)};
}

template <class T>
std::vector<std::vector<T>> host_cons(eb::HipExecSpace const& ex, eb::EulerConsArrays<view_t<T>> const& a)
{
    return {eb::to_host(ex, a.d), eb::to_host(ex, a.e), eb::to_host(ex, a.mx0), eb::to_host(ex, a.mx1),
            eb::to_host(ex, a.mx2)};
}

template <class T>
bool same(std::vector<std::vector<T>> const& a, std::vector<std::vector<T>> const& b)
{
    for (int f = 0; f < 5; ++f) {
        if (a[f].size() != b[f].size() || std::memcmp(a[f].data(), b[f].data(), a[f].size() * sizeof(T)) != 0) {
            return false;
        }
    }
    return true;
}

template <class T>
void reference_types_checks(char const* tag)
{
    eb::HipExecSpace const ex;
    int const n0 = 37, n1 = 20, n2 = 17;
    State<T> a(ex, n0, n1, n2), b(ex, n0, n1, n2);
    eb::init_random(ex, a.prim, 99);
    eb::init_random(ex, b.prim, 99);
    T const g = T(1.4);
    eb::PerfectGas<T> const eos(g);
    perfect_gas_again<T> const eos2(g);
    eb::UniformMesh3d<T> const mesh(T(1) / 35, T(1) / 18, T(1) / 15);
    mesh_again<T> const mesh2(T(1) / 35, T(1) / 18, T(1) / 15);
    eb::prim_to_cons(ex, eb::as_const(a.prim), a.cons, eos);
    eb::prim_to_cons(ex, eb::as_const(b.prim), b.cons, eos2);
    expect(same(host_cons(ex, a.cons), host_cons(ex, b.cons)), std::string(tag) + " prim_to_cons: plug-in == library");
    T const dt1 = eb::time_step(ex, eb::as_const(a.prim), eos, mesh);
    T const dt2 = eb::time_step(ex, eb::as_const(b.prim), eos2, mesh2);
    expect(dt1 == dt2, std::string(tag) + " time_step: plug-in == library");
    eb::godunov(ex, eb::as_const(a.prim), a.cons, eos, mesh, eb::hllc {}, T(0.5) * dt1);
    eb::godunov(ex, eb::as_const(b.prim), b.cons, eos2, mesh2, hllc_again {}, T(0.5) * dt2);
    expect(same(host_cons(ex, a.cons), host_cons(ex, b.cons)), std::string(tag) + " godunov: plug-in == library");
    eb::cons_to_prim(ex, eb::as_const(a.cons), a.prim, eos);
    eb::cons_to_prim(ex, eb::as_const(b.cons), b.prim, eos2);
    expect(same(host_prim(ex, a.prim), host_prim(ex, b.prim)), std::string(tag) + " cons_to_prim: plug-in == library");
}

// Host evaluation of the reference's loops (godunov.hpp:40-90, cons_to_prim.hpp:28-31,
// prim_to_cons.hpp:28-31, time_step.hpp:18-39) through the given functors.
template <class T, class Solver, class EOS, class Mesh>
struct HostReference
{
    int n0, n1, n2;
    Solver rs;
    EOS eos;
    Mesh mesh;

    std::size_t at(int i, int j, int k) const { return std::size_t(i) + std::size_t(n0) * (j + std::size_t(n1) * k); }

    eb::EulerPrim<T> q(std::vector<std::vector<T>> const& p, std::size_t c) const
    {
        return {p[0][c], p[1][c], p[2][c], p[3][c], p[4][c]};
    }

    void prim_to_cons(std::vector<std::vector<T>> const& p, std::vector<std::vector<T>>& u) const
    {
        EB_IEEE_ORDER
        for (std::size_t c = 0; c < p[0].size(); ++c) {
            eb::EulerCons<T> const v = eb::to_cons(q(p, c), eos.internal_energy(p[0][c], p[1][c]));
            u[0][c] = v.d, u[1][c] = v.e, u[2][c] = v.mx0, u[3][c] = v.mx1, u[4][c] = v.mx2;
        }
    }

    void cons_to_prim(std::vector<std::vector<T>> const& u, std::vector<std::vector<T>>& p) const
    {
        EB_IEEE_ORDER
        for (std::size_t c = 0; c < u[0].size(); ++c) {
            eb::EulerCons<T> const v {u[0][c], u[1][c], u[2][c], u[3][c], u[4][c]};
            eb::EulerPrim<T> const w = eb::to_prim(v, eos.pressure(v.d, eb::internal_energy(v)));
            p[0][c] = w.d, p[1][c] = w.p, p[2][c] = w.ux0, p[3][c] = w.ux1, p[4][c] = w.ux2;
        }
    }

    T time_step(std::vector<std::vector<T>> const& p) const
    {
        EB_IEEE_ORDER
        T const i0 = 1 / mesh.dx0(), i1 = 1 / mesh.dx1(), i2 = 1 / mesh.dx2();
        T dt = std::numeric_limits<T>::max();
        for (std::size_t c = 0; c < p[0].size(); ++c) {
            T const cs = eos.speed_of_sound(p[0][c], p[1][c]);
            T const invdt = ((cs + std::fabs(p[2][c])) * i0) + ((cs + std::fabs(p[3][c])) * i1)
                            + ((cs + std::fabs(p[4][c])) * i2);
            T const cand = T(1) / invdt;
            dt = cand < dt ? cand : dt;
        }
        return dt;
    }

    void godunov(std::vector<std::vector<T>> const& p, std::vector<std::vector<T>>& u, T dt) const
    {
        EB_IEEE_ORDER
        T const ds[3] = {mesh.ds0(), mesh.ds1(), mesh.ds2()};
        T const dtodv = dt / mesh.dv();
        for (int k = 1; k < n2 - 1; ++k) {
            for (int j = 1; j < n1 - 1; ++j) {
                for (int i = 1; i < n0 - 1; ++i) {
                    eb::EulerPrim<T> const c = q(p, at(i, j, k));
                    eb::EulerFlux<T> fl {};
                    auto add = [&](T s, eb::EulerFlux<T> const& r, eb::EulerFlux<T> const& l) {
                        fl.d += s * (r.d - l.d);
                        fl.e += s * (r.e - l.e);
                        fl.mx0 += s * (r.mx0 - l.mx0);
                        fl.mx1 += s * (r.mx1 - l.mx1);
                        fl.mx2 += s * (r.mx2 - l.mx2);
                    };
                    using d0 = std::integral_constant<std::size_t, 0>;
                    using d1 = std::integral_constant<std::size_t, 1>;
                    using d2 = std::integral_constant<std::size_t, 2>;
                    add(ds[0], rs(d0 {}, eos, c, q(p, at(i + 1, j, k))), rs(d0 {}, eos, q(p, at(i - 1, j, k)), c));
                    add(ds[1], rs(d1 {}, eos, c, q(p, at(i, j + 1, k))), rs(d1 {}, eos, q(p, at(i, j - 1, k)), c));
                    add(ds[2], rs(d2 {}, eos, c, q(p, at(i, j, k + 1))), rs(d2 {}, eos, q(p, at(i, j, k - 1)), c));
                    std::size_t const x = at(i, j, k);
                    u[0][x] -= dtodv * fl.d;
                    u[1][x] -= dtodv * fl.e;
                    u[2][x] -= dtodv * fl.mx0;
                    u[3][x] -= dtodv * fl.mx1;
                    u[4][x] -= dtodv * fl.mx2;
                }
            }
        }
    }
};

template <class T>
void caller_types_checks(char const* tag)
{
    eb::HipExecSpace const ex;
    int const n0 = 23, n1 = 15, n2 = 12;
    State<T> s(ex, n0, n1, n2);
    eb::init_random(ex, s.prim, 7);
    stiffened_gas<T> const eos {T(4.4), T(0.6)};
    box_mesh<T> const mesh {T(0.05), T(0.0625), T(0.08)};
    rusanov const rs;
    HostReference<T, rusanov, stiffened_gas<T>, box_mesh<T>> const ref {n0, n1, n2, rs, eos, mesh};

    auto p = host_prim(ex, s.prim);
    auto u = p;
    eb::prim_to_cons(ex, eb::as_const(s.prim), s.cons, eos);
    ref.prim_to_cons(p, u);
    expect(same(host_cons(ex, s.cons), u), std::string(tag) + " stiffened-gas prim_to_cons on the GPU == host");

    T const dt = eb::time_step(ex, eb::as_const(s.prim), eos, mesh);
    expect(dt == ref.time_step(p) && dt > 0, std::string(tag) + " stiffened-gas time_step on the GPU == host");

    auto const before = u;
    eb::godunov(ex, eb::as_const(s.prim), s.cons, eos, mesh, rs, T(0.4) * dt);
    ref.godunov(p, u, T(0.4) * dt);
    auto const got = host_cons(ex, s.cons);
    std::size_t changed = 0;
    for (std::size_t c = 0; c < got[1].size(); ++c) {
        changed += got[1][c] != before[1][c] ? 1 : 0;
    }
    expect(same(got, u) && changed > got[1].size() / 2,
           std::string(tag) + " Rusanov + stiffened gas + caller mesh godunov on the GPU == host");
    eb::cons_to_prim(ex, eb::as_const(s.cons), s.prim, eos);
    ref.cons_to_prim(u, p);
    expect(same(host_prim(ex, s.prim), p), std::string(tag) + " stiffened-gas cons_to_prim on the GPU == host");
}

} // namespace

int main()
{
    try {
        reference_types_checks<double>("f64");
        reference_types_checks<float>("f32");
        caller_types_checks<double>("f64");
        caller_types_checks<float>("f32");
    } catch (std::exception const& e) {
        std::cout << "FAIL exception: " << e.what() << '\n';
        return 1;
    }
    std::cout << (g_failures == 0 ? "ALL PASS" : "FAILURES") << '\n';
    return g_failures == 0 ? 0 : 1;
}
