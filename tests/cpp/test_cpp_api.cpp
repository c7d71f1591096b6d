// Checks of the C++ operator API over the HIP library (run by tests/test_gpu_cpp_api.py):
// view helpers, closed forms (uniform state: godunov is the identity and
// time_step = 1/(3c/dx)), the reference-order loop against the fused step loop
// (eager and graph-replayed, bitwise), real/index type instantiations and the
// error behaviour. Prints one line per check; exit status 0 iff all pass.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <euler/euler.hpp>

namespace {

int g_failures = 0;

void expect(bool ok, std::string const& what)
{
    std::cout << (ok ? "PASS " : "FAIL ") << what << '\n' << std::flush;
    if (!ok) {
        ++g_failures;
    }
}

template <class T, class I>
using view_t = eb::mdspan<T, eb::dextents<I, 3>, eb::layout_left>;

template <class A>
bool same_bits(eb::HipExecSpace const& ex, A const& a, A const& b)
{
    auto eq = [&](auto const& x, auto const& y) {
        auto const hx = eb::to_host(ex, x);
        auto const hy = eb::to_host(ex, y);
        return std::memcmp(hx.data(), hy.data(), hx.size() * sizeof(hx[0])) == 0;
    };
    return eq(a.d, b.d) && eq(a.ux0, b.ux0) && eq(a.ux1, b.ux1) && eq(a.ux2, b.ux2);
}

template <class T, class I>
void uniform_state_checks(std::string const& tag)
{
    eb::HipExecSpace const ex;
    I const n = 10;
    auto const pa = eb::create_prim_arrays_1d<T>(ex, std::size_t(n) * n * n);
    auto const ca = eb::create_cons_arrays_1d<T>(ex, std::size_t(n) * n * n);
    auto const prim = eb::to_mdspan<view_t<T, I>>(pa, n, n, n);
    auto const cons = eb::to_mdspan<view_t<T, I>>(ca, n, n, n);
    expect(eb::size(prim) == std::size_t(1000) && eb::size_bytes(prim) == 5 * sizeof(T) * 1000,
           tag + " size / size_bytes");
    std::vector<T> one(1000, T(1)), zero(1000, T(0));
    eb::from_host(ex, one, prim.d);
    eb::from_host(ex, one, prim.p);
    eb::from_host(ex, zero, prim.ux0);
    eb::from_host(ex, zero, prim.ux1);
    eb::from_host(ex, zero, prim.ux2);
    eb::PerfectGas<T> const eos(T(1.4));
    eb::UniformMesh3d<T> const mesh(T(1), T(1), T(1));
    eb::prim_to_cons(ex, eb::as_const(prim), cons, eos);
    T const dt = eb::time_step(ex, eb::as_const(prim), eos, mesh);
    T const c = eos.speed_of_sound(T(1), T(1));
    T const expect_dt = T(1) / ((c * T(1)) + (c * T(1)) + (c * T(1)));
    expect(dt == expect_dt, tag + " time_step of a uniform state is 1/(3c)");
    auto const before = eb::to_host(ex, cons.e);
    eb::godunov(ex, eb::as_const(prim), cons, eos, mesh, eb::hllc {}, T(0.5) * dt);
    auto const after = eb::to_host(ex, cons.e);
    expect(std::memcmp(before.data(), after.data(), before.size() * sizeof(T)) == 0,
           tag + " godunov on a uniform state is the identity");
    T const e_expect = eos.internal_energy(T(1), T(1));
    expect(after[555] == e_expect, tag + " prim_to_cons energy = p/(gamma-1) at rest");
}

template <class T, class I>
void loop_checks(std::string const& tag)
{
    eb::HipExecSpace const ex;
    I const na = 18;
    T const dx = T(1) / T(16);
    eb::PerfectGas<T> const eos(T(1.4));
    eb::UniformMesh3d<T> const mesh(dx, dx, dx);
    std::size_t const ncell = std::size_t(na) * na * na;
    auto const pa = eb::create_prim_arrays_1d<T>(ex, ncell);
    auto const ca = eb::create_cons_arrays_1d<T>(ex, ncell);
    auto const prim = eb::to_mdspan<view_t<T, I>>(pa, na, na, na);
    auto const cons = eb::to_mdspan<view_t<T, I>>(ca, na, na, na);
    eb::init_implode(ex, prim, mesh);
    eb::prim_to_cons(ex, eb::as_const(prim), cons, eos);
    // second copy for the fused loops
    auto const pb = eb::create_prim_arrays_1d<T>(ex, ncell);
    auto const cb = eb::create_cons_arrays_1d<T>(ex, ncell);
    auto const prim2 = eb::to_mdspan<view_t<T, I>>(pb, na, na, na);
    auto const cons2 = eb::to_mdspan<view_t<T, I>>(cb, na, na, na);
    auto const pc = eb::create_prim_arrays_1d<T>(ex, ncell);
    auto const cc = eb::create_cons_arrays_1d<T>(ex, ncell);
    auto const prim3 = eb::to_mdspan<view_t<T, I>>(pc, na, na, na);
    auto const cons3 = eb::to_mdspan<view_t<T, I>>(cc, na, na, na);
    for (auto const* p : {&prim2, &prim3}) {
        eb::init_implode(ex, *p, mesh);
    }
    eb::prim_to_cons(ex, eb::as_const(prim2), cons2, eos);
    eb::prim_to_cons(ex, eb::as_const(prim3), cons3, eos);

    T const cfl = T(0.49);
    long const nt = 40; // the graph-replayed run covers one 32-step graph, pairs and an odd step
    double t_ref = 0;
    for (long it = 0; it < nt; ++it) {
        T const dt = eb::time_step(ex, eb::as_const(prim), eos, mesh);
        eb::godunov(ex, eb::as_const(prim), cons, eos, mesh, eb::hllc {}, cfl * dt);
        eb::cons_to_prim(ex, eb::as_const(cons), prim, eos);
        t_ref += static_cast<double>(cfl * dt);
    }
    {
        eb::StepLoop<T, I> loop(ex, prim2, cons2, eos, mesh, cfl);
        loop.run(nt, false);
        double const t = loop.time();
        loop.finish();
        expect(same_bits(ex, prim, prim2), tag + " fused loop == reference-order loop (prims, bitwise)");
        expect(same_bits(ex, eb::EulerPrimArrays<view_t<T, I>> {cons.d, cons.e, cons.mx0, cons.mx1, cons.mx2},
                         eb::EulerPrimArrays<view_t<T, I>> {cons2.d, cons2.e, cons2.mx0, cons2.mx1, cons2.mx2}),
               tag + " fused loop == reference-order loop (cons, bitwise)");
        expect(std::abs(t - t_ref) <= 1e-12 * t_ref, tag + " simulated time accumulated on the device");
    }
    {
        eb::StepLoop<T, I> loop(ex, prim3, cons3, eos, mesh, cfl);
        loop.run(3, false);
        loop.run(nt - 3, true);
        loop.finish();
        expect(same_bits(ex, prim, prim3), tag + " graph-replayed loop == reference-order loop (bitwise)");
    }
}

void error_checks()
{
    eb::HipExecSpace const ex;
    auto const pa = eb::create_prim_arrays_1d<double>(ex, 64);
    bool threw = false;
    try {
        (void)eb::to_mdspan<view_t<double, int>>(pa, 5, 5, 5);
    } catch (eb::Error const& e) {
        threw = e.status() == EB_ERR_INVALID_ARGUMENT;
    }
    expect(threw, "to_mdspan beyond the allocation throws");

    // 2^31 cells do not fit a 32-bit index: rejected before any memory is touched.
    double* const none = nullptr;
    view_t<double, int> big(none + 0, 2048, 1024, 1024);
    eb::EulerPrimArrays<view_t<double, int>> const bigp {big, big, big, big, big};
    eb::EulerConsArrays<view_t<double, int>> const bigc {big, big, big, big, big};
    int status = EB_OK;
    try {
        eb::prim_to_cons(ex, eb::as_const(bigp), bigc, eb::PerfectGas<double>(1.4));
    } catch (eb::Error const& e) {
        status = e.status();
    }
    expect(status == EB_ERR_INDEX_OVERFLOW || status == EB_ERR_INVALID_ARGUMENT,
           "32-bit index overflow is an error, not a launch");

    auto const pb = eb::create_prim_arrays_1d<double>(ex, 1000);
    auto const cb = eb::create_cons_arrays_1d<double>(ex, 1000);
    auto const p10 = eb::to_mdspan<view_t<double, int>>(pb, 10, 10, 10);
    auto const c8 = eb::to_mdspan<view_t<double, int>>(cb, 8, 8, 8);
    threw = false;
    try {
        eb::prim_to_cons(ex, eb::as_const(p10), c8, eb::PerfectGas<double>(1.4));
    } catch (eb::Error const& e) {
        threw = e.status() == EB_ERR_INVALID_ARGUMENT;
    }
    expect(threw, "prim / cons extents mismatch throws");
}

void npy_check()
{
    std::vector<double> v(2 * 3 * 4);
    for (std::size_t i = 0; i < v.size(); ++i) {
        v[i] = static_cast<double>(i) + 0.5;
    }
    std::ostringstream os;
    eb::save_npy(os, v.data(), {2, 3, 4});
    std::string const s = os.str();
    bool ok = s.size() == 128 + v.size() * 8 && s.compare(1, 5, "NUMPY") == 0 && s[127] == '\n'
              && s.find("'fortran_order': True") != std::string::npos
              && s.find("'shape': (2, 3, 4, ), }") != std::string::npos;
    double last = 0;
    std::memcpy(&last, s.data() + s.size() - 8, 8);
    expect(ok && last == 23.5, "save_npy header and payload");
}

// benchmark_utils.hpp counterparts: int_cast range checks, init_from_value and
// init_from_state fill every element (ghosts included) with the value.
void harness_checks()
{
    expect(eb::int_cast<int>(std::int64_t(2147483647)) == 2147483647, "int_cast keeps an in-range value");
    bool threw = false;
    try {
        (void)eb::int_cast<int>(std::int64_t(2147483648LL));
    } catch (std::runtime_error const&) {
        threw = true;
    }
    expect(threw, "int_cast throws past the target range");
    threw = false;
    try {
        (void)eb::int_cast<unsigned>(-1);
    } catch (std::runtime_error const&) {
        threw = true;
    }
    expect(threw, "int_cast throws on a negative value for an unsigned target");

    eb::HipExecSpace const ex;
    auto const pa = eb::create_prim_arrays_1d<float>(ex, 5 * 6 * 7);
    auto const prim = eb::to_mdspan<view_t<float, int>>(pa, 5, 6, 7);
    auto const ca = eb::create_cons_arrays_1d<float>(ex, 5 * 6 * 7);
    auto const cons = eb::to_mdspan<view_t<float, int>>(ca, 5, 6, 7);
    eb::init_from_state(ex, prim, eb::EulerPrim<float> {1.5f, 2.5f, -0.25f, 0.5f, 3.0f});
    eb::init_from_state(ex, cons, eb::EulerCons<float> {1.f, 2.5f, 0.f, -1.f, 7.f});
    ex.fence();
    bool ok = true;
    float const pv[5] = {1.5f, 2.5f, -0.25f, 0.5f, 3.0f};
    float const cv[5] = {1.f, 2.5f, 0.f, -1.f, 7.f};
    auto const pf = std::array {prim.d, prim.p, prim.ux0, prim.ux1, prim.ux2};
    auto const cf = std::array {cons.d, cons.e, cons.mx0, cons.mx1, cons.mx2};
    for (int f = 0; f < 5; ++f) {
        for (float x : eb::to_host(ex, pf[f])) {
            ok = ok && x == pv[f];
        }
        for (float x : eb::to_host(ex, cf[f])) {
            ok = ok && x == cv[f];
        }
    }
    expect(ok, "init_from_state fills all 210 cells of every field");
}

} // namespace

int main()
{
    try {
        uniform_state_checks<double, int>("f64/int");
        uniform_state_checks<float, unsigned>("f32/unsigned");
        uniform_state_checks<double, std::int64_t>("f64/int64");
        loop_checks<double, int>("f64/int");
        loop_checks<float, std::int64_t>("f32/int64");
        error_checks();
        npy_check();
        harness_checks();
    } catch (std::exception const& e) {
        std::cout << "FAIL exception: " << e.what() << '\n';
        return 1;
    }
    std::cout << (g_failures == 0 ? "ALL PASS" : "FAILURES") << '\n';
    return g_failures == 0 ? 0 : 1;
}
