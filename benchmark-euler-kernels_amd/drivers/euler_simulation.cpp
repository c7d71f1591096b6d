// Step-loop driver (the counterpart of simulations/euler_simulation.cpp): an
// implosion on the unit cube, nx^3 interior cells, nt steps of
// dt = time_step; godunov(cfl * dt); cons_to_prim, pressure written as .npy
// every output_freq steps, and the cell-update rate printed at the end.
//
//   euler_simulation [--nx 128] [--nt 200] [--output-freq 10] [--cfl 0.49]
//                    [--gamma 1.4] [--real f64|f32] [--index int|unsigned|long]
//                    [--loop reference|fused|graph] [--init implode|diagonal]
//                    [--out DIR]
//
// --loop reference runs the reference's three operators per step with the dt
// round trip to the host; fused runs one fused launch per step with dt on the
// device; graph replays pairs of fused steps from a HIP graph. All three give
// bitwise identical states.
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include <euler/euler.hpp>

namespace {

struct Options
{
    long nx = 128;
    long nt = 200;
    long output_freq = 10;
    double cfl = 0.49;
    double gamma = 1.4;
    std::string real = "f64";
    std::string index = "int";
    std::string loop = "reference";
    std::string init = "implode";
    std::filesystem::path out = ".";
};

Options parse(int argc, char** argv)
{
    Options o;
    for (int a = 1; a < argc; ++a) {
        std::string const key = argv[a];
        if (a + 1 >= argc) {
            throw std::invalid_argument("missing value for " + key);
        }
        std::string const v = argv[++a];
        if (key == "--nx") {
            o.nx = std::stol(v);
        } else if (key == "--nt") {
            o.nt = std::stol(v);
        } else if (key == "--output-freq") {
            o.output_freq = std::stol(v);
        } else if (key == "--cfl") {
            o.cfl = std::stod(v);
        } else if (key == "--gamma") {
            o.gamma = std::stod(v);
        } else if (key == "--real") {
            o.real = v;
        } else if (key == "--index") {
            o.index = v;
        } else if (key == "--loop") {
            o.loop = v;
        } else if (key == "--init") {
            o.init = v;
        } else if (key == "--out") {
            o.out = v;
        } else {
            throw std::invalid_argument("unknown option " + key);
        }
    }
    if (o.nx < 1 || o.nt < 0) {
        throw std::invalid_argument("nx must be >= 1 and nt >= 0");
    }
    return o;
}

template <class T, class IndexType>
int run(Options const& o)
{
    using view = eb::mdspan<T, eb::dextents<IndexType, 3>, eb::layout_left>;
    IndexType const na = static_cast<IndexType>(o.nx + 2);
    T const dx = T(1) / static_cast<T>(o.nx);
    eb::PerfectGas<T> const eos(static_cast<T>(o.gamma));
    eb::UniformMesh3d<T> const mesh(dx, dx, dx);
    eb::hllc const riemann_solver;
    eb::DefaultExecutionSpace const exec_space;
    std::size_t const ncell = static_cast<std::size_t>(na) * na * na;
    auto const prims_alloc = eb::create_prim_arrays_1d<T>(exec_space, ncell);
    auto const prim_arrays = eb::to_mdspan<view>(prims_alloc, na, na, na);
    auto const cons_alloc = eb::create_cons_arrays_1d<T>(exec_space, ncell);
    auto const cons_arrays = eb::to_mdspan<view>(cons_alloc, na, na, na);

    if (o.init == "implode") {
        eb::init_implode(exec_space, prim_arrays, mesh);
    } else if (o.init == "diagonal") {
        eb::init_diagonal(exec_space, prim_arrays, mesh);
    } else {
        throw std::invalid_argument("unknown init " + o.init);
    }
    eb::prim_to_cons(exec_space, eb::as_const(prim_arrays), cons_arrays, eos);

    auto save = [&](long it, auto const& prims) {
        if (o.output_freq > 0 && it % o.output_freq == 0) {
            std::stringstream ss;
            ss << "test_" << std::setfill('0') << std::setw(10) << it << ".npy";
            eb::save_npy(o.out / ss.str(), exec_space, prims.p);
            std::cout << "Saving " << ss.str() << " done\n" << std::flush;
        }
    };

    exec_space.fence();
    double time_in_seconds = 0;
    double sim_time = 0;
    if (o.loop == "reference") {
        auto const start = std::chrono::steady_clock::now();
        for (long it = 1; it <= o.nt; ++it) {
            T const dt = eb::time_step(exec_space, eb::as_const(prim_arrays), eos, mesh);
            eb::godunov(exec_space, eb::as_const(prim_arrays), cons_arrays, eos, mesh, riemann_solver,
                        static_cast<T>(o.cfl) * dt);
            eb::cons_to_prim(exec_space, eb::as_const(cons_arrays), prim_arrays, eos);
            sim_time += static_cast<double>(static_cast<T>(o.cfl) * dt);
            save(it, prim_arrays);
        }
        exec_space.fence();
        time_in_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    } else if (o.loop == "fused" || o.loop == "graph") {
        eb::StepLoop<T, IndexType> loop(exec_space, prim_arrays, cons_arrays, eos, mesh, static_cast<T>(o.cfl));
        bool const graph = o.loop == "graph";
        long const chunk = o.output_freq > 0 ? o.output_freq : o.nt;
        auto const start = std::chrono::steady_clock::now();
        for (long it = 0; it < o.nt;) {
            long const n = std::min(chunk, o.nt - it);
            loop.run(n, graph);
            it += n;
            save(it, loop.current_prim());
        }
        exec_space.fence();
        time_in_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        sim_time = loop.time();
        loop.finish();
    } else {
        throw std::invalid_argument("unknown loop " + o.loop);
    }
    double const cells_updated = static_cast<double>(o.nt) * static_cast<double>(o.nx * o.nx * o.nx);
    std::cout << std::setprecision(17) << "simulated_time " << sim_time << '\n';
    std::cout << cells_updated / time_in_seconds * 1e-6 << '\n';
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    try {
        Options const o = parse(argc, argv);
        if (o.real == "f64" && o.index == "int") {
            return run<double, int>(o);
        }
        if (o.real == "f64" && o.index == "long") {
            return run<double, std::int64_t>(o);
        }
        if (o.real == "f32" && o.index == "int") {
            return run<float, int>(o);
        }
        if (o.real == "f32" && o.index == "long") {
            return run<float, std::int64_t>(o);
        }
        if (o.real == "f64" && o.index == "unsigned") {
            return run<double, unsigned>(o);
        }
        if (o.real == "f32" && o.index == "unsigned") {
            return run<float, unsigned>(o);
        }
        throw std::invalid_argument("--real f64|f32, --index int|unsigned|long");
    } catch (std::exception const& e) {
        std::cerr << "euler_simulation: " << e.what() << '\n';
        return 1;
    }
}
