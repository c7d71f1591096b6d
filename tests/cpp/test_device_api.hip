// Device-side checks of the C++ API (run by tests/test_gpu_cpp_api.py), built with
// hipcc's default contraction (-ffp-contract=fast-honor-pragmas, which fuses
// multiply-adds wherever no pragma forbids it), so that the point functions' own
// switch (EB_IEEE_ORDER) is what keeps their arithmetic in the reference's order:
//   (default)  the public hllc functor, get<Dir>, load / store through raw
//              pointers, mdspans and subspan views, all called inside kernels,
//              against the library's HLLC kernel (eb_hllc_flux_*, bitwise equal
//              to the oracle) and against the same functor on the host;
//   --large    int and unsigned index types past the 32-bit kernels' limit
//              (1290^3 cells with int, 1300^3 with unsigned, FP32): the C++
//              operators route them to the 64-bit kernels, and cells beyond
//              2^31 are checked against host evaluations of the reference's
//              formulas through the same public point functions.
// One line per check, exit status 0 iff all pass.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <random>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>

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

void hip_ok(hipError_t e, char const* what)
{
    if (e != hipSuccess) {
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }
}

template <std::size_t D>
using dir_t = std::integral_constant<std::size_t, D>;

// One face per thread: fluxes of qL | qR across a face normal to dir through the
// public functor, inputs and outputs moved with load / store.
template <class T>
__global__ void k_functor_flux(int dir, eb::EulerPrimArrays<T const*> qL, eb::EulerPrimArrays<T const*> qR,
                               eb::EulerConsArrays<T*> out, int n, T gamma)
{
    int const f = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (f >= n) {
        return;
    }
    eb::PerfectGas<T> const eos(gamma);
    eb::EulerPrim<T> const a = eb::load(qL, f);
    eb::EulerPrim<T> const b = eb::load(qR, f);
    eb::hllc const solver;
    eb::EulerFlux<T> fl {};
    if (dir == 0) {
        fl = solver(dir_t<0> {}, eos, a, b);
    } else if (dir == 1) {
        fl = solver(dir_t<1> {}, eos, a, b);
    } else {
        fl = solver(dir_t<2> {}, eos, a, b);
    }
    // get<Dir> of the flux is its normal momentum component
    T const normal = dir == 0 ? eb::get(dir_t<0> {}, fl) : (dir == 1 ? eb::get(dir_t<1> {}, fl) : eb::get(dir_t<2> {}, fl));
    T const check = dir == 0 ? fl.mx0 : (dir == 1 ? fl.mx1 : fl.mx2);
    eb::store(eb::EulerCons<T> {fl.d, fl.e, fl.mx0, fl.mx1, fl.mx2}, out, f);
    if (!(normal == check || normal != normal)) {
        out.d[f] = T(-12345); // flagged for the host
    }
}

// out_box(i, j, k) = in_box(i, j, k) for the cells of two layout_stride subspans.
template <class View, class ViewC>
__global__ void k_copy_box(eb::EulerPrimArrays<ViewC> in, eb::EulerPrimArrays<View> out)
{
    using I = typename View::index_type;
    I const i = static_cast<I>(threadIdx.x);
    I const j = static_cast<I>(blockIdx.x);
    I const k = static_cast<I>(blockIdx.y);
    if (i < in.d.extent(0) && j < in.d.extent(1) && k < in.d.extent(2)) {
        eb::store(eb::load(in, i, j, k), out, i, j, k);
    }
}

template <class T>
struct Faces
{
    std::vector<int> dir;
    std::vector<T> q[2][5]; // side, field
};

template <class T>
Faces<T> make_faces(int n)
{
    std::mt19937_64 rng(4242);
    std::uniform_real_distribution<double> pos(0.1, 3.0);
    std::normal_distribution<double> vel(0.0, 1.2);
    Faces<T> fc;
    fc.dir.resize(n);
    for (auto& side : fc.q) {
        for (auto& v : side) {
            v.resize(n);
        }
    }
    for (int f = 0; f < n; ++f) {
        int const d = f % 3;
        fc.dir[f] = d;
        double q[2][5];
        for (auto& s : q) {
            s[0] = pos(rng);
            s[1] = pos(rng);
            s[2] = vel(rng);
            s[3] = vel(rng);
            s[4] = vel(rng);
        }
        switch ((f / 3) % 5) {
        case 1: // supersonic, alternating sign
            q[0][2 + d] = (f % 2 ? 1 : -1) * (5.0 + std::fabs(q[0][2 + d]));
            q[1][2 + d] = (f % 2 ? 1 : -1) * (5.0 + std::fabs(q[1][2 + d]));
            break;
        case 2: // mirror states: u* = 0
            for (int c = 0; c < 5; ++c) {
                q[1][c] = q[0][c];
            }
            q[1][2 + d] = -q[0][2 + d];
            break;
        case 3: // min/max ties
            q[1][0] = q[0][0];
            q[1][1] = q[0][1];
            q[1][2 + d] = q[0][2 + d];
            break;
        default:
            break;
        }
        for (int s = 0; s < 2; ++s) {
            for (int c = 0; c < 5; ++c) {
                fc.q[s][c][f] = static_cast<T>(q[s][c]);
            }
        }
    }
    return fc;
}

template <class T>
struct DevBuf
{
    T* p = nullptr;
    explicit DevBuf(std::size_t n)
    {
        hip_ok(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)), "hipMalloc");
    }
    ~DevBuf()
    {
        (void)hipFree(p);
    }
    DevBuf(DevBuf const&) = delete;
    DevBuf& operator=(DevBuf const&) = delete;
};

template <class T>
struct capi_flux;
template <>
struct capi_flux<double>
{
    static constexpr auto fn = &eb_hllc_flux_f64;
    using prim = eb_prim_f64;
    using cons = eb_cons_f64;
};
template <>
struct capi_flux<float>
{
    static constexpr auto fn = &eb_hllc_flux_f32;
    using prim = eb_prim_f32;
    using cons = eb_cons_f32;
};

template <class T>
void functor_checks(char const* tag)
{
    int const n = 3000;
    T const gamma = static_cast<T>(1.4);
    Faces<T> const fc = make_faces<T>(n);
    // device copies: [side][field], plus outputs of the functor and of the library
    std::vector<DevBuf<T>*> bufs;
    auto dev = [&](std::vector<T> const& h) {
        bufs.push_back(new DevBuf<T>(h.size()));
        hip_ok(hipMemcpy(bufs.back()->p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        return bufs.back()->p;
    };
    T* qd[2][5];
    for (int s = 0; s < 2; ++s) {
        for (int c = 0; c < 5; ++c) {
            qd[s][c] = dev(fc.q[s][c]);
        }
    }
    std::vector<T> zero(n, T(0));
    T* fo[5];
    T* lo[5];
    for (int c = 0; c < 5; ++c) {
        fo[c] = dev(zero);
        lo[c] = dev(zero);
    }
    eb::EulerPrimArrays<T const*> const L {qd[0][0], qd[0][1], qd[0][2], qd[0][3], qd[0][4]};
    eb::EulerPrimArrays<T const*> const R {qd[1][0], qd[1][1], qd[1][2], qd[1][3], qd[1][4]};
    eb::EulerConsArrays<T*> const out {fo[0], fo[1], fo[2], fo[3], fo[4]};
    // face f is normal to direction f % 3: three launches (one per direction) over
    // all faces, and per face the result of the launch of its own direction
    std::vector<T> got[5];
    std::vector<T> lib[5];
    for (int c = 0; c < 5; ++c) {
        got[c].assign(n, T(0));
        lib[c].assign(n, T(0));
    }
    typename capi_flux<T>::prim const pl {qd[0][0], qd[0][1], qd[0][2], qd[0][3], qd[0][4]};
    typename capi_flux<T>::prim const pr {qd[1][0], qd[1][1], qd[1][2], qd[1][3], qd[1][4]};
    typename capi_flux<T>::cons const lc {lo[0], lo[1], lo[2], lo[3], lo[4]};
    std::vector<T> tmp(n);
    for (int d = 0; d < 3; ++d) {
        hipLaunchKernelGGL(k_functor_flux<T>, dim3((n + 255) / 256), dim3(256), 0, 0, d, L, R, out, n, gamma);
        hip_ok(hipGetLastError(), "launch");
        if (capi_flux<T>::fn(nullptr, d, &pl, &pr, &lc, n, gamma) != EB_OK) {
            throw std::runtime_error(std::string("eb_hllc_flux: ") + eb_last_error());
        }
        hip_ok(hipDeviceSynchronize(), "sync");
        for (int c = 0; c < 5; ++c) {
            hip_ok(hipMemcpy(tmp.data(), fo[c], n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
            for (int f = d; f < n; f += 3) {
                got[c][f] = tmp[f];
            }
            hip_ok(hipMemcpy(tmp.data(), lo[c], n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
            for (int f = d; f < n; f += 3) {
                lib[c][f] = tmp[f];
            }
        }
    }
    auto same = [](T a, T b) { return std::memcmp(&a, &b, sizeof(T)) == 0 || (a != a && b != b); };
    int mism_lib = 0;
    int mism_host = 0;
    eb::PerfectGas<T> const eos(gamma);
    eb::hllc const solver;
    for (int f = 0; f < n; ++f) {
        eb::EulerPrim<T> const a {fc.q[0][0][f], fc.q[0][1][f], fc.q[0][2][f], fc.q[0][3][f], fc.q[0][4][f]};
        eb::EulerPrim<T> const b {fc.q[1][0][f], fc.q[1][1][f], fc.q[1][2][f], fc.q[1][3][f], fc.q[1][4][f]};
        eb::EulerFlux<T> h {};
        switch (fc.dir[f]) {
        case 0: h = solver(dir_t<0> {}, eos, a, b); break;
        case 1: h = solver(dir_t<1> {}, eos, a, b); break;
        default: h = solver(dir_t<2> {}, eos, a, b); break;
        }
        T const hv[5] = {h.d, h.e, h.mx0, h.mx1, h.mx2};
        for (int c = 0; c < 5; ++c) {
            mism_lib += same(got[c][f], lib[c][f]) ? 0 : 1;
            mism_host += same(got[c][f], hv[c]) ? 0 : 1;
        }
    }
    expect(mism_lib == 0, std::string(tag) + " hllc functor in a kernel == library HLLC kernel, bitwise (3000 faces)");
    expect(mism_host == 0, std::string(tag) + " hllc functor in a kernel == the functor on the host, bitwise");
    for (auto* b : bufs) {
        delete b;
    }
}

void view_checks()
{
    using view = eb::mdspan<double, eb::dextents<int, 3>, eb::layout_left>;
    int const n0 = 37, n1 = 11, n2 = 9;
    std::size_t const nc = std::size_t(n0) * n1 * n2;
    eb::HipExecSpace const ex;
    auto const ia = eb::create_prim_arrays_1d<double>(ex, nc);
    auto const oa = eb::create_prim_arrays_1d<double>(ex, nc);
    auto const in = eb::to_mdspan<view>(ia, n0, n1, n2);
    auto const outv = eb::to_mdspan<view>(oa, n0, n1, n2);
    std::vector<double> h(nc);
    auto const fields_in = std::array {in.d, in.p, in.ux0, in.ux1, in.ux2};
    for (int f = 0; f < 5; ++f) {
        for (std::size_t q = 0; q < nc; ++q) {
            h[q] = 1000.0 * f + static_cast<double>(q);
        }
        eb::from_host(ex, h, fields_in[f]);
    }
    // interior box of `in` into the box shifted by (-1, 0, +1)... kept in bounds:
    // source [1, n0-1) x [1, n1-1) x [1, n2-1), destination [0, n0-2) x [1, n1-1) x [2, n2)
    auto const src = eb::subspan(eb::as_const(in), std::pair {1, n0 - 1}, std::pair {1, n1 - 1}, std::pair {1, n2 - 1});
    auto const dst = eb::subspan(outv, std::pair {0, n0 - 2}, std::pair {1, n1 - 1}, std::pair {2, n2});
    using SV = std::remove_cv_t<decltype(dst.d)>;
    using SC = std::remove_cv_t<decltype(src.d)>;
    static_assert(std::is_same_v<typename SV::layout_type, eb::layout_stride>);
    hipLaunchKernelGGL((k_copy_box<SV, SC>), dim3(n1 - 2, n2 - 2), dim3(64), 0, ex.stream(), src, dst);
    hip_ok(hipGetLastError(), "launch");
    ex.fence();
    auto const fields_out = std::array {outv.d, outv.p, outv.ux0, outv.ux1, outv.ux2};
    bool ok = true;
    for (int f = 0; f < 5; ++f) {
        auto const o = eb::to_host(ex, fields_out[f]);
        for (int k = 0; k < n2; ++k) {
            for (int j = 0; j < n1; ++j) {
                for (int i = 0; i < n0; ++i) {
                    bool const inside = i < n0 - 2 && j >= 1 && j < n1 - 1 && k >= 2;
                    double const want = inside ? 1000.0 * f + ((i + 1) + n0 * (j + n1 * (k - 1))) : 0.0;
                    ok = ok && o[i + n0 * (j + n1 * k)] == want;
                }
            }
        }
    }
    expect(ok, "load / store through layout_stride subspans inside a kernel move exactly the box");
}

// ------------------------------------------------------------------ large ---

// Host restatement of one interior cell's update (godunov.hpp:48-90) through the
// public functor, for the uniform state with one planted cell.
template <class T, class Q>
eb::EulerCons<T> host_update(Q const& q, eb::EulerCons<T> u, eb::PerfectGas<T> const& eos,
                             eb::UniformMesh3d<T> const& mesh, T dt, long i, long j, long k)
{
#pragma clang fp contract(off)
    eb::hllc const s;
    T const ds[3] = {mesh.ds0(), mesh.ds1(), mesh.ds2()};
    eb::EulerFlux<T> acc {};
    auto add = [&](T dsd, eb::EulerFlux<T> const& r, eb::EulerFlux<T> const& l) {
        acc.d += dsd * (r.d - l.d);
        acc.e += dsd * (r.e - l.e);
        acc.mx0 += dsd * (r.mx0 - l.mx0);
        acc.mx1 += dsd * (r.mx1 - l.mx1);
        acc.mx2 += dsd * (r.mx2 - l.mx2);
    };
    add(ds[0], s(dir_t<0> {}, eos, q(i, j, k), q(i + 1, j, k)), s(dir_t<0> {}, eos, q(i - 1, j, k), q(i, j, k)));
    add(ds[1], s(dir_t<1> {}, eos, q(i, j, k), q(i, j + 1, k)), s(dir_t<1> {}, eos, q(i, j - 1, k), q(i, j, k)));
    add(ds[2], s(dir_t<2> {}, eos, q(i, j, k), q(i, j, k + 1)), s(dir_t<2> {}, eos, q(i, j, k - 1), q(i, j, k)));
    T const dtodv = dt / mesh.dv();
    u.d -= dtodv * acc.d;
    u.e -= dtodv * acc.e;
    u.mx0 -= dtodv * acc.mx0;
    u.mx1 -= dtodv * acc.mx1;
    u.mx2 -= dtodv * acc.mx2;
    return u;
}

template <class I>
void large_checks(char const* tag, long n)
{
    using T = float;
    using view = eb::mdspan<T, eb::dextents<I, 3>, eb::layout_left>;
    long const ncell = n * n * n;
    std::cout << "large " << tag << ": " << n << "^3 = " << ncell << " cells, 32-bit kernels take at most "
              << EB_MAX_CELLS_INDEX32 << ", index width " << eb::detail::index_bits<I>(n, n, n) << '\n'
              << std::flush;
    eb::HipExecSpace const ex;
    auto const pa = eb::create_prim_arrays_1d<T>(ex, ncell);
    auto const ca = eb::create_cons_arrays_1d<T>(ex, ncell);
    auto const prim = eb::to_mdspan<view>(pa, static_cast<I>(n), static_cast<I>(n), static_cast<I>(n));
    auto const cons = eb::to_mdspan<view>(ca, static_cast<I>(n), static_cast<I>(n), static_cast<I>(n));
    eb::EulerPrim<T> const bg {1.0f, 1.0f, 0.1f, -0.05f, 0.02f};
    eb::init_from_state(ex, prim, bg);
    // planted cell near the far corner, its linear index above 2^31
    long const pi = n - 3, pj = n - 2, pk = n - 2;
    long const lin = pi + n * (pj + n * pk);
    eb::EulerPrim<T> const hot {0.5f, 3.0f, 2.5f, -1.5f, 0.75f};
    auto const fields = std::array {prim.d, prim.p, prim.ux0, prim.ux1, prim.ux2};
    T const hv[5] = {hot.d, hot.p, hot.ux0, hot.ux1, hot.ux2};
    for (int f = 0; f < 5; ++f) {
        hip_ok(hipMemcpyAsync(fields[f].data_handle() + lin, &hv[f], sizeof(T), hipMemcpyHostToDevice, ex.stream()),
               "hipMemcpy");
    }
    ex.fence();
    eb::PerfectGas<T> const eos(1.4f);
    T const dx = 1.0f / static_cast<T>(n - 2);
    eb::UniformMesh3d<T> const mesh(dx, dx, dx);

    // prim_to_cons over every cell, read back around the planted cell
    eb::prim_to_cons(ex, eb::as_const(prim), cons, eos);
    // ex's stream does not synchronise with the null stream of hipMemcpy: fence first
    auto read_cons = [&](long i, long j, long k) {
        ex.fence();
        long const q = i + n * (j + n * k);
        T v[5];
        auto const cf = std::array {cons.d, cons.e, cons.mx0, cons.mx1, cons.mx2};
        for (int f = 0; f < 5; ++f) {
            hip_ok(hipMemcpy(&v[f], cf[f].data_handle() + q, sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        }
        return eb::EulerCons<T> {v[0], v[1], v[2], v[3], v[4]};
    };
    auto read_prim = [&](long i, long j, long k) {
        ex.fence();
        long const q = i + n * (j + n * k);
        T v[5];
        for (int f = 0; f < 5; ++f) {
            hip_ok(hipMemcpy(&v[f], fields[f].data_handle() + q, sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        }
        return eb::EulerPrim<T> {v[0], v[1], v[2], v[3], v[4]};
    };
    auto same = [](eb::EulerCons<T> const& a, eb::EulerCons<T> const& b) {
        return std::memcmp(&a, &b, sizeof(a)) == 0;
    };
    eb::EulerCons<T> const hot_c = eb::to_cons(hot, eos.internal_energy(hot.d, hot.p));
    eb::EulerCons<T> const bg_c = eb::to_cons(bg, eos.internal_energy(bg.d, bg.p));
    expect(same(read_cons(pi, pj, pk), hot_c) && same(read_cons(pi + 1, pj, pk), bg_c)
                   && same(read_cons(n - 1, n - 1, n - 1), bg_c),
           std::string(tag) + " prim_to_cons at cells past 2^31 (planted and last cell)");

    // time_step: the planted cell is the fastest, so the minimum is its dt
    T const dt = eb::time_step(ex, eb::as_const(prim), eos, mesh);
    auto inv = [&](eb::EulerPrim<T> const& q) {
#pragma clang fp contract(off)
        T const c = eos.speed_of_sound(q.d, q.p);
        return ((c + std::fabs(q.ux0)) * (1.0f / dx)) + ((c + std::fabs(q.ux1)) * (1.0f / dx))
               + ((c + std::fabs(q.ux2)) * (1.0f / dx));
    };
    T const dt_hot = 1.0f / inv(hot);
    expect(dt == dt_hot && dt_hot < 1.0f / inv(bg), std::string(tag) + " time_step finds the cell past 2^31");

    // godunov: the planted cell and its six neighbours against the host update
    T const step = 0.5f * dt;
    eb::godunov(ex, eb::as_const(prim), cons, eos, mesh, eb::hllc {}, step);
    auto q = [&](long i, long j, long k) { return (i == pi && j == pj && k == pk) ? hot : bg; };
    bool ok = true;
    long const nb[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (auto const& o : nb) {
        long const i = pi + o[0], j = pj + o[1], k = pk + o[2];
        if (i < 1 || j < 1 || k < 1 || i > n - 2 || j > n - 2 || k > n - 2) {
            ok = ok && same(read_cons(i, j, k), bg_c); // ghost cell: never updated
            continue;
        }
        eb::EulerCons<T> const u0 = (o[0] == 0 && o[1] == 0 && o[2] == 0) ? hot_c : bg_c;
        ok = ok && same(read_cons(i, j, k), host_update(q, u0, eos, mesh, step, i, j, k));
    }
    expect(ok, std::string(tag) + " godunov at the planted cell past 2^31 and its neighbours, bitwise");

    // cons_to_prim back: the planted cell's primitives from its updated conservatives
    eb::cons_to_prim(ex, eb::as_const(cons), prim, eos);
    eb::EulerCons<T> const uh = read_cons(pi, pj, pk);
    eb::EulerPrim<T> const want = eb::to_prim(uh, eos.pressure(uh.d, eb::internal_energy(uh)));
    eb::EulerPrim<T> const gotp = read_prim(pi, pj, pk);
    expect(std::memcmp(&want, &gotp, sizeof(want)) == 0, std::string(tag) + " cons_to_prim at the cell past 2^31");
}

} // namespace

int main(int argc, char** argv)
{
    try {
        if (argc > 1 && std::string(argv[1]) == "--large") {
            large_checks<int>("int", 1290);
            large_checks<unsigned>("unsigned", 1300);
        } else {
            functor_checks<double>("f64");
            functor_checks<float>("f32");
            view_checks();
        }
    } catch (std::exception const& e) {
        std::cout << "FAIL exception: " << e.what() << '\n';
        return 1;
    }
    std::cout << (g_failures == 0 ? "ALL PASS" : "FAILURES") << '\n';
    return g_failures == 0 ? 0 : 1;
}
