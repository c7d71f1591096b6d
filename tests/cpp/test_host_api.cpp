// Host-side checks of the C++ API that need no GPU (run by tests/test_host_api.py):
//   checks          views (mdspan, submdspan, subspan), load/store, get<Dir>, the
//                   index width chosen per index type, the .npy header bytes;
//                   prints one line per check, exit status 0 iff all pass.
//   hllc f64|f32 IN OUT N
//                   reads N faces (dir, qL[5], qR[5], eleven reals each) from IN,
//                   writes the hllc functor's N fluxes (five reals each) to OUT,
//                   for a bitwise comparison with the oracle.
// Built with -ffp-contract=off, so the functor's arithmetic is plain IEEE.
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
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

template <class I>
using view3 = eb::mdspan<double, eb::dextents<I, 3>, eb::layout_left>;

void view_checks()
{
    int const n0 = 6, n1 = 5, n2 = 4;
    int const nc = n0 * n1 * n2;
    std::vector<double> buf(5 * nc);
    for (std::size_t q = 0; q < buf.size(); ++q) {
        buf[q] = static_cast<double>(q);
    }
    auto field = [&](int f) { return view3<int>(buf.data() + f * nc, n0, n1, n2); };
    eb::EulerPrimArrays<view3<int>> const prim {field(0), field(1), field(2), field(3), field(4)};

    bool ok = true;
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                ok = ok && &prim.d(i, j, k) == buf.data() + (i + n0 * (j + n1 * k));
            }
        }
    }
    expect(ok && prim.d.stride(0) == 1 && prim.d.stride(1) == n0 && prim.d.stride(2) == n0 * n1,
           "layout_left maps (i, j, k) to i + n0*(j + n1*k)");

    // interior box: ranges on every rank -> layout_stride, same strides
    auto const inner = eb::subspan(prim, std::pair {1, n0 - 1}, std::pair {1, n1 - 1}, std::pair {1, n2 - 1});
    using inner_t = std::remove_cv_t<decltype(inner.d)>;
    static_assert(std::is_same_v<inner_t::layout_type, eb::layout_stride>);
    static_assert(inner_t::rank() == 3);
    ok = inner.d.extent(0) == n0 - 2 && inner.d.extent(1) == n1 - 2 && inner.d.extent(2) == n2 - 2;
    for (int k = 0; k < n2 - 2; ++k) {
        for (int j = 0; j < n1 - 2; ++j) {
            for (int i = 0; i < n0 - 2; ++i) {
                ok = ok && &inner.ux1(i, j, k) == &prim.ux1(i + 1, j + 1, k + 1)
                     && &inner.d(i, j, k) == &prim.d(i + 1, j + 1, k + 1);
            }
        }
    }
    expect(ok, "subspan of ranges: interior box, layout_stride, every field");

    // a k-plane: full, full, index -> layout_left of rank 2
    auto const plane = eb::subspan(prim, eb::full_extent, eb::full_extent, 3);
    using plane_t = std::remove_cv_t<decltype(plane.p)>;
    static_assert(std::is_same_v<plane_t::layout_type, eb::layout_left>);
    static_assert(plane_t::rank() == 2);
    ok = plane.p.extent(0) == n0 && plane.p.extent(1) == n1;
    for (int j = 0; j < n1; ++j) {
        for (int i = 0; i < n0; ++i) {
            ok = ok && &plane.p(i, j) == &prim.p(i, j, 3);
        }
    }
    expect(ok, "subspan full, full, index: a plane in layout_left");

    // x-range of row j = 2 across k: range, index, full -> layout_stride of rank 2
    auto const rows = eb::subspan(prim, std::tuple {2, 5}, 2, eb::full_extent);
    using rows_t = std::remove_cv_t<decltype(rows.d)>;
    static_assert(std::is_same_v<rows_t::layout_type, eb::layout_stride>);
    static_assert(rows_t::rank() == 2);
    ok = rows.d.extent(0) == 3 && rows.d.extent(1) == n2 && rows.d.stride(1) == n0 * n1;
    for (int k = 0; k < n2; ++k) {
        for (int i = 0; i < 3; ++i) {
            ok = ok && &rows.ux2(i, k) == &prim.ux2(2 + i, 2, k);
        }
    }
    expect(ok, "subspan range, index, full: strided rows");

    // an x-line: range, index, index -> layout_left of rank 1
    auto const line = eb::submdspan(prim.d, std::pair {1, 4}, 3, 2);
    using line_t = std::remove_cv_t<decltype(line)>;
    static_assert(std::is_same_v<line_t::layout_type, eb::layout_left>);
    expect(line.extent(0) == 3 && &line(0) == &prim.d(1, 3, 2) && &line(2) == &prim.d(3, 3, 2),
           "submdspan range, index, index: a unit-stride line");

    // every rank an index -> rank 0, the one element
    auto const cell = eb::submdspan(prim.ux0, 4, 1, 2);
    static_assert(std::remove_cv_t<decltype(cell)>::rank() == 0);
    expect(&cell() == &prim.ux0(4, 1, 2) && cell.size() == 1, "submdspan of indices only: one element");

    // load / store through views (also through a subspan) and through raw pointers
    auto const q = eb::load(prim, 2, 3, 1);
    ok = q.d == prim.d(2, 3, 1) && q.p == prim.p(2, 3, 1) && q.ux0 == prim.ux0(2, 3, 1)
         && q.ux1 == prim.ux1(2, 3, 1) && q.ux2 == prim.ux2(2, 3, 1);
    eb::store(eb::EulerPrim<double> {-1, -2, -3, -4, -5}, inner, 0, 1, 1);
    auto const back = eb::load(prim, 1, 2, 2);
    ok = ok && back.d == -1 && back.p == -2 && back.ux0 == -3 && back.ux1 == -4 && back.ux2 == -5;
    auto const ptrs = eb::data_handle(prim);
    int const lin = 1 + n0 * (2 + n1 * 2);
    auto const viap = eb::load(ptrs, lin);
    ok = ok && viap.ux1 == -4;
    eb::store(eb::EulerPrim<double> {7, 8, 9, 10, 11}, ptrs, lin);
    ok = ok && prim.d(1, 2, 2) == 7 && prim.ux2(1, 2, 2) == 11;
    eb::EulerConsArrays<view3<int>> const cons {field(0), field(1), field(2), field(3), field(4)};
    eb::store(eb::EulerCons<double> {1.5, 2.5, 3.5, 4.5, 5.5}, cons, 5, 4, 3);
    auto const u = eb::load(eb::as_const(cons), 5, 4, 3);
    ok = ok && u.d == 1.5 && u.e == 2.5 && u.mx0 == 3.5 && u.mx1 == 4.5 && u.mx2 == 5.5;
    expect(ok, "load / store of EulerPrim and EulerCons through views, subspans and pointers");

    // get<Dir>: the normal component of each point struct
    eb::EulerPrim<double> const pp {1, 2, 3, 4, 5};
    eb::EulerCons<double> const cc {1, 2, 6, 7, 8};
    eb::EulerFlux<double> const ff {1, 2, 9, 10, 11};
    using d0 = std::integral_constant<std::size_t, 0>;
    using d1 = std::integral_constant<std::size_t, 1>;
    using d2 = std::integral_constant<std::size_t, 2>;
    expect(eb::get(d0 {}, pp) == 3 && eb::get(d1 {}, pp) == 4 && eb::get(d2 {}, pp) == 5 && eb::get(d0 {}, cc) == 6
                   && eb::get(d2 {}, cc) == 8 && eb::get(d1 {}, ff) == 10 && eb::get(d2 {}, ff) == 11,
           "get<Dir> of prim, cons and flux");
}

// Index width of the kernels per index type: 32 up to EB_MAX_CELLS_INDEX32 cells,
// 64 above it while the count fits the type, else 32 (which the C ABI rejects).
void index_width_checks()
{
    using eb::detail::index_bits;
    bool ok = index_bits<int>(514, 514, 514) == 32 && index_bits<unsigned>(514, 514, 514) == 32
              && index_bits<std::int64_t>(4, 4, 4) == 64;
    // 1290^3 = 2,146,689,000: fits int, above the 32-bit kernels' limit
    ok = ok && index_bits<int>(1290, 1290, 1290) == 64;
    // 2^31 cells: past int, within unsigned
    ok = ok && index_bits<int>(2048, 1024, 1024) == 32 && index_bits<unsigned>(2048, 1024, 1024) == 64;
    // 1300^3 = 2,197,000,000 and 65535 * 65535 = 4,294,836,225 < 2^32: unsigned
    ok = ok && index_bits<unsigned>(1300, 1300, 1300) == 64 && index_bits<unsigned>(65535, 65535, 1) == 64;
    // 2^33 cells: past unsigned
    ok = ok && index_bits<unsigned>(65536, 65536, 2) == 32;
    ok = ok && index_bits<std::uint16_t>(40, 40, 40) == 32;
    expect(ok, "index width: int and unsigned reach their full ranges through the 64-bit kernels");
}

// The reference writer's header (simulations/save_npy.cpp:34-64), restated
// independently: dict with each extent followed by ", ", padded with spaces so
// that 10 + dict + '\n' is a multiple of 64.
std::string reference_npy_header(char const* descr, std::array<std::size_t, 3> const& shape)
{
    std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': True, 'shape': (";
    for (std::size_t e : shape) {
        dict += std::to_string(e) + ", ";
    }
    dict += "), }";
    std::size_t const total = 10 + dict.size() + 1;
    dict += std::string((total + 63) / 64 * 64 - total, ' ') + "\n";
    std::string head = std::string("\x93NUMPY", 6) + '\x01' + '\x00';
    head += static_cast<char>(dict.size() & 0xff);
    head += static_cast<char>(dict.size() >> 8);
    return head + dict;
}

void npy_checks()
{
    bool ok = true;
    for (auto const& shape : {std::array<std::size_t, 3> {2, 3, 4}, std::array<std::size_t, 3> {130, 130, 130},
                              std::array<std::size_t, 3> {1, 1, 1}, std::array<std::size_t, 3> {1000000, 7, 123456}}) {
        std::size_t const n = shape[0] == 1000000 ? 0 : shape[0] * shape[1] * shape[2];
        std::vector<double> v(n);
        for (std::size_t i = 0; i < n; ++i) {
            v[i] = static_cast<double>(i) + 0.5;
        }
        std::ostringstream os;
        if (n == 0) {
            // header only: stream a shape whose payload is not written here
            std::ostringstream hdr;
            std::array<std::size_t, 3> const z {shape[0], shape[1], 0};
            eb::save_npy(hdr, v.data(), z);
            ok = ok && hdr.str() == reference_npy_header("<f8", z);
            continue;
        }
        eb::save_npy(os, v.data(), shape);
        std::string const s = os.str();
        std::string const h = reference_npy_header("<f8", shape);
        ok = ok && s.size() == h.size() + n * 8 && s.compare(0, h.size(), h) == 0
             && std::memcmp(s.data() + h.size(), v.data(), n * 8) == 0;
        std::vector<float> w(n, 1.25f);
        std::ostringstream os4;
        eb::save_npy(os4, w.data(), shape);
        std::string const h4 = reference_npy_header("<f4", shape);
        ok = ok && os4.str().compare(0, h4.size(), h4) == 0 && os4.str().size() == h4.size() + n * 4;
    }
    std::ostringstream os;
    std::vector<double> v(24);
    eb::save_npy(os, v.data(), {2, 3, 4});
    std::string const expected_dict = "{'descr': '<f8', 'fortran_order': True, 'shape': (2, 3, 4, ), }";
    ok = ok && os.str().compare(10, expected_dict.size(), expected_dict) == 0 && os.str().size() == 128 + 24 * 8
         && os.str()[127] == '\n';
    expect(ok, "save_npy header bytes equal the reference writer's");
}

template <class T>
int run_hllc(char const* in_path, char const* out_path, long n)
{
    std::vector<T> in(static_cast<std::size_t>(n) * 11);
    std::ifstream is(in_path, std::ios::binary);
    is.read(reinterpret_cast<char*>(in.data()), static_cast<std::streamsize>(in.size() * sizeof(T)));
    if (!is) {
        std::cerr << "cannot read " << in_path << '\n';
        return 2;
    }
    eb::PerfectGas<T> const eos(static_cast<T>(1.4));
    eb::hllc const solver;
    std::vector<T> out(static_cast<std::size_t>(n) * 5);
    for (long f = 0; f < n; ++f) {
        T const* r = in.data() + f * 11;
        eb::EulerPrim<T> const qL {r[1], r[2], r[3], r[4], r[5]};
        eb::EulerPrim<T> const qR {r[6], r[7], r[8], r[9], r[10]};
        eb::EulerFlux<T> fl {};
        switch (static_cast<int>(r[0])) {
        case 0: fl = solver(std::integral_constant<std::size_t, 0> {}, eos, qL, qR); break;
        case 1: fl = solver(std::integral_constant<std::size_t, 1> {}, eos, qL, qR); break;
        default: fl = solver(std::integral_constant<std::size_t, 2> {}, eos, qL, qR); break;
        }
        T* o = out.data() + f * 5;
        o[0] = fl.d;
        o[1] = fl.e;
        o[2] = fl.mx0;
        o[3] = fl.mx1;
        o[4] = fl.mx2;
    }
    std::ofstream os(out_path, std::ios::binary);
    os.write(reinterpret_cast<char const*>(out.data()), static_cast<std::streamsize>(out.size() * sizeof(T)));
    return os ? 0 : 2;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc == 6 && std::string(argv[1]) == "hllc") {
        long const n = std::stol(argv[5]);
        return std::string(argv[2]) == "f32" ? run_hllc<float>(argv[3], argv[4], n)
                                              : run_hllc<double>(argv[3], argv[4], n);
    }
    view_checks();
    index_width_checks();
    npy_checks();
    std::cout << (g_failures == 0 ? "ALL PASS" : "FAILURES") << '\n';
    return g_failures == 0 ? 0 : 1;
}
