// Operator benchmarks with Google-Benchmark-shaped output: the counterpart of the
// reference's benchmarks/ executable (benchmark_{godunov,time_step,cons_to_prim,
// prim_to_cons,euler_simulation}.cpp + benchmark_main.cpp), so that tools reading
// the reference's JSON (plot.py) read these results unchanged.
//
//   euler_benchmarks [--benchmark_filter=REGEX] [--benchmark_min_time=0.5s]
//                    [--benchmark_out=FILE] [--benchmark_out_format=json]
//                    [--real=f64|f32] [--index=int|long]
//
// Each benchmark sets up the state the reference's does (uniform states, the
// implosion for EulerSimulation), fences after every iteration as the reference
// loop does, and reports per-iteration real and CPU time in ns plus the
// reference's counters: cells_per_second and bytes_per_second are iteration-
// invariant rates over CPU time (benchmark_utils.cpp:5-15). The host thread spins
// while it waits for the GPU, so CPU time and real time agree.
//
// Iterations grow until a run lasts at least the minimum time, as Google Benchmark
// does. The real type and index type are run-time options here (the reference
// fixes them at compile time, benchmarks/real_type.hpp and index_type.hpp; its
// defaults are double and int).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <regex>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include <euler/benchmark_utils.hpp>
#include <euler/euler.hpp>

namespace {

struct Options
{
    std::string filter = ".*";
    double min_time = 0.5;
    std::string out;
    std::string out_format = "json";
    std::string real = "f64";
    std::string index = "int";
};

double parse_seconds(std::string v)
{
    if (!v.empty() && v.back() == 's') {
        v.pop_back();
    }
    double const t = std::stod(v);
    if (!(t > 0)) {
        throw std::invalid_argument("--benchmark_min_time must be positive");
    }
    return t;
}

Options parse(int argc, char** argv)
{
    Options o;
    for (int a = 1; a < argc; ++a) {
        std::string const arg = argv[a];
        auto const eq = arg.find('=');
        if (arg.rfind("--", 0) != 0 || eq == std::string::npos) {
            throw std::invalid_argument("options are --name=value, got " + arg);
        }
        std::string const key = arg.substr(0, eq);
        std::string const v = arg.substr(eq + 1);
        if (key == "--benchmark_filter") {
            o.filter = v;
        } else if (key == "--benchmark_min_time") {
            o.min_time = parse_seconds(v);
        } else if (key == "--benchmark_out") {
            o.out = v;
        } else if (key == "--benchmark_out_format") {
            o.out_format = v;
        } else if (key == "--real") {
            o.real = v;
        } else if (key == "--index") {
            o.index = v;
        } else {
            throw std::invalid_argument("unknown option " + key);
        }
    }
    if (o.out_format != "json") {
        throw std::invalid_argument("--benchmark_out_format: only json is written");
    }
    return o;
}

double cpu_seconds()
{
    timespec ts {};
    clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &ts);
    return static_cast<double>(ts.tv_sec) + 1e-9 * static_cast<double>(ts.tv_nsec);
}

struct Result
{
    std::string name;
    std::int64_t iterations = 0;
    double real_ns = 0; // per iteration
    double cpu_ns = 0;  // per iteration
    double cells = 0;   // per iteration, iteration-invariant counter (0: not set)
    double bytes = 0;
};

// A benchmark instance: `setup` prepares the state and returns the iteration body
// and the per-iteration cell and byte counts (set_constant_*_processed).
struct Case
{
    std::string name;
    std::function<void(std::function<void()>&, double&, double&)> setup;
};

Result run_case(Case const& c, double min_time)
{
    std::function<void()> body;
    double cells = 0;
    double bytes = 0;
    c.setup(body, cells, bytes);
    body(); // first touch outside the timing, as a warm-up iteration
    std::int64_t iters = 1;
    for (;;) {
        double const c0 = cpu_seconds();
        auto const r0 = std::chrono::steady_clock::now();
        for (std::int64_t i = 0; i < iters; ++i) {
            body();
        }
        double const real = std::chrono::duration<double>(std::chrono::steady_clock::now() - r0).count();
        double const cpu = cpu_seconds() - c0;
        if (real >= min_time || iters >= 1000000000) {
            Result r;
            r.name = c.name;
            r.iterations = iters;
            r.real_ns = real / static_cast<double>(iters) * 1e9;
            r.cpu_ns = cpu / static_cast<double>(iters) * 1e9;
            r.cells = cells;
            r.bytes = bytes;
            return r;
        }
        // grow like Google Benchmark: aim 40% past the minimum, at most 10x per round
        double mult = real > 0 ? min_time * 1.4 / real : 10.0;
        mult = mult > 10.0 ? 10.0 : (mult < 1.0 ? 2.0 : mult);
        std::int64_t const next = static_cast<std::int64_t>(static_cast<double>(iters) * mult) + 1;
        iters = next > iters ? next : iters + 1;
    }
}

template <class T, class IndexType>
std::vector<Case> cases(eb::HipExecSpace const& ex)
{
    using view = eb::mdspan<T, eb::dextents<IndexType, 3>, eb::layout_left>;
    std::vector<Case> out;
    std::vector<long> dense;
    for (long n = 8; n <= 31; n += 8) { // DenseRange(8, 31, 8)
        dense.push_back(n);
    }
    for (long n = 32; n <= 320; n += 32) { // DenseRange(32, 320, 32)
        dense.push_back(n);
    }

    // benchmark_godunov.cpp:17-54: uniform state, unit cells, dt = 1e-9
    for (long r : dense) {
        out.push_back({"Godunov/" + std::to_string(r), [&ex, r](std::function<void()>& body, double& cells,
                                                                double& bytes) {
                           auto const n = eb::int_cast<IndexType>(r + 2);
                           eb::PerfectGas<T> const eos(1.4);
                           eb::UniformMesh3d<T> const mesh(1., 1., 1.);
                           T const dt = 1E-9;
                           std::size_t const cnt = static_cast<std::size_t>(n) * n * n;
                           auto const pa = eb::create_prim_arrays_1d<T>(ex, cnt);
                           auto const prim = eb::to_mdspan<view>(pa, n, n, n);
                           auto const ca = eb::create_cons_arrays_1d<T>(ex, cnt);
                           auto const cons = eb::to_mdspan<view>(ca, n, n, n);
                           eb::EulerPrim<T> const q {1, 1, 0, 0, 0};
                           eb::init_from_state(ex, prim, q);
                           eb::init_from_state(ex, cons, eb::to_cons(q, eos.internal_energy(q.d, q.p)));
                           ex.fence();
                           body = [=, &ex]() {
                               (void)pa;
                               (void)ca;
                               eb::godunov(ex, eb::as_const(prim), cons, eos, mesh, eb::hllc(), dt);
                               ex.fence();
                           };
                           cells = static_cast<double>(eb::size(cons));
                           bytes = static_cast<double>(eb::size_bytes(prim) + 2 * eb::size_bytes(cons));
                       }});
    }
    // benchmark_time_step.cpp:17-37: uniform state, dt returned to the host
    for (long r : dense) {
        out.push_back({"TimeStep/" + std::to_string(r), [&ex, r](std::function<void()>& body, double& cells,
                                                                 double& bytes) {
                           auto const n = eb::int_cast<IndexType>(r);
                           eb::PerfectGas<T> const eos(1.4);
                           eb::UniformMesh3d<T> const mesh(1., 1., 1.);
                           std::size_t const cnt = static_cast<std::size_t>(n) * n * n;
                           auto const pa = eb::create_prim_arrays_1d<T>(ex, cnt);
                           auto const prim = eb::to_mdspan<view>(pa, n, n, n);
                           eb::init_from_state(ex, prim, eb::EulerPrim<T> {1, 1, 0, 0, 0});
                           ex.fence();
                           body = [=, &ex]() {
                               (void)pa;
                               volatile T dt = eb::time_step(ex, eb::as_const(prim), eos, mesh);
                               (void)dt;
                           };
                           cells = static_cast<double>(eb::size(prim));
                           bytes = static_cast<double>(eb::size_bytes(prim));
                       }});
    }
    // benchmark_cons_to_prim.cpp:15-41: uniform conservative state
    for (long r : dense) {
        out.push_back({"ConsToPrim/" + std::to_string(r), [&ex, r](std::function<void()>& body, double& cells,
                                                                   double& bytes) {
                           auto const n = eb::int_cast<IndexType>(r);
                           eb::PerfectGas<T> const eos(1.4);
                           std::size_t const cnt = static_cast<std::size_t>(n) * n * n;
                           auto const pa = eb::create_prim_arrays_1d<T>(ex, cnt);
                           auto const prim = eb::to_mdspan<view>(pa, n, n, n);
                           auto const ca = eb::create_cons_arrays_1d<T>(ex, cnt);
                           auto const cons = eb::to_mdspan<view>(ca, n, n, n);
                           eb::init_from_state(ex, cons, eb::EulerCons<T> {1, T(1 / 0.4), 0, 0, 0});
                           ex.fence();
                           body = [=, &ex]() {
                               (void)pa;
                               (void)ca;
                               eb::cons_to_prim(ex, eb::as_const(cons), prim, eos);
                               ex.fence();
                           };
                           cells = static_cast<double>(eb::size(prim));
                           bytes = static_cast<double>(eb::size_bytes(prim) + eb::size_bytes(cons));
                       }});
    }
    // benchmark_prim_to_cons.cpp:15-41: uniform primitive state
    for (long r : dense) {
        out.push_back({"PrimToCons/" + std::to_string(r), [&ex, r](std::function<void()>& body, double& cells,
                                                                   double& bytes) {
                           auto const n = eb::int_cast<IndexType>(r);
                           eb::PerfectGas<T> const eos(1.4);
                           std::size_t const cnt = static_cast<std::size_t>(n) * n * n;
                           auto const pa = eb::create_prim_arrays_1d<T>(ex, cnt);
                           auto const prim = eb::to_mdspan<view>(pa, n, n, n);
                           auto const ca = eb::create_cons_arrays_1d<T>(ex, cnt);
                           auto const cons = eb::to_mdspan<view>(ca, n, n, n);
                           eb::init_from_state(ex, prim, eb::EulerPrim<T> {1, 1, 0, 0, 0});
                           ex.fence();
                           body = [=, &ex]() {
                               (void)pa;
                               (void)ca;
                               eb::prim_to_cons(ex, eb::as_const(prim), cons, eos);
                               ex.fence();
                           };
                           cells = static_cast<double>(eb::size(cons));
                           bytes = static_cast<double>(eb::size_bytes(prim) + eb::size_bytes(cons));
                       }});
    }
    // benchmark_euler_simulation.cpp:20-62: implosion, one reference step per
    // iteration (time_step with its host round trip, godunov, cons_to_prim)
    for (long r = 16; r <= 320; r += 32) { // DenseRange(16, 320, 32)
        out.push_back({"EulerSimulation/" + std::to_string(r), [&ex, r](std::function<void()>& body,
                                                                        double& cells, double& bytes) {
                           auto const nx = eb::int_cast<IndexType>(r);
                           T const cfl_factor = 0.49;
                           T const dx = T(1.) / static_cast<T>(nx);
                           eb::PerfectGas<T> const eos(1.4);
                           eb::UniformMesh3d<T> const mesh(dx, dx, dx);
                           IndexType const na = nx + 2;
                           std::size_t const cnt = static_cast<std::size_t>(na) * na * na;
                           auto const pa = eb::create_prim_arrays_1d<T>(ex, cnt);
                           auto const prim = eb::to_mdspan<view>(pa, na, na, na);
                           auto const ca = eb::create_cons_arrays_1d<T>(ex, cnt);
                           auto const cons = eb::to_mdspan<view>(ca, na, na, na);
                           eb::init_implode(ex, prim, mesh);
                           eb::prim_to_cons(ex, eb::as_const(prim), cons, eos);
                           ex.fence();
                           body = [=, &ex]() {
                               (void)pa;
                               (void)ca;
                               T const dt = eb::time_step(ex, eb::as_const(prim), eos, mesh);
                               eb::godunov(ex, eb::as_const(prim), cons, eos, mesh, eb::hllc(), cfl_factor * dt);
                               eb::cons_to_prim(ex, eb::as_const(cons), prim, eos);
                               ex.fence();
                           };
                           cells = static_cast<double>(eb::size(cons));
                           bytes = 0; // the reference sets no byte counter here
                       }});
    }
    return out;
}

std::string json_escape(std::string const& s)
{
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') {
            o += '\\';
        }
        o += c;
    }
    return o;
}

void write_json(std::ostream& os, std::vector<Result> const& res, Options const& o, char const* exe)
{
    char host[256] = {0};
    gethostname(host, sizeof(host) - 1);
    std::time_t const now = std::time(nullptr);
    char date[64];
    std::strftime(date, sizeof(date), "%Y-%m-%dT%H:%M:%S%z", std::localtime(&now));
    hipDeviceProp_t prop {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipGetDeviceProperties(&prop, dev);
    os << std::setprecision(17);
    os << "{\n  \"context\": {\n";
    os << "    \"date\": \"" << date << "\",\n";
    os << "    \"host_name\": \"" << json_escape(host) << "\",\n";
    os << "    \"executable\": \"" << json_escape(exe) << "\",\n";
    os << "    \"num_cpus\": " << sysconf(_SC_NPROCESSORS_ONLN) << ",\n";
    os << "    \"mhz_per_cpu\": 0,\n";
    os << "    \"cpu_scaling_enabled\": false,\n";
    os << "    \"library_build_type\": \"release\",\n";
    std::string const dev_name = prop.name[0] != '\0' ? std::string(prop.name) + " " : std::string();
    os << "    \"device\": \"" << json_escape(dev_name) << json_escape(prop.gcnArchName) << "\",\n";
    os << "    \"real_type\": \"" << (o.real == "f64" ? "double" : "float") << "\",\n";
    os << "    \"index_type\": \"" << (o.index == "int" ? "int" : "long") << "\",\n";
    os << "    \"rate_time_base\": \"cpu_time (the host spins while it waits for the GPU)\"\n";
    os << "  },\n  \"benchmarks\": [\n";
    for (std::size_t i = 0; i < res.size(); ++i) {
        Result const& r = res[i];
        double const cpu_s = r.cpu_ns * 1e-9;
        os << "    {\n";
        os << "      \"name\": \"" << r.name << "\",\n";
        os << "      \"family_index\": 0,\n";
        os << "      \"per_family_instance_index\": " << i << ",\n";
        os << "      \"run_name\": \"" << r.name << "\",\n";
        os << "      \"run_type\": \"iteration\",\n";
        os << "      \"repetitions\": 1,\n";
        os << "      \"repetition_index\": 0,\n";
        os << "      \"threads\": 1,\n";
        os << "      \"iterations\": " << r.iterations << ",\n";
        os << "      \"real_time\": " << r.real_ns << ",\n";
        os << "      \"cpu_time\": " << r.cpu_ns << ",\n";
        os << "      \"time_unit\": \"ns\"";
        if (r.bytes > 0) {
            os << ",\n      \"bytes_per_second\": " << r.bytes / cpu_s;
        }
        if (r.cells > 0) {
            os << ",\n      \"cells_per_second\": " << r.cells / cpu_s;
        }
        os << "\n    }" << (i + 1 < res.size() ? "," : "") << "\n";
    }
    os << "  ]\n}\n";
}

template <class T, class IndexType>
int run(Options const& o, char const* exe)
{
    eb::DefaultExecutionSpace const ex;
    std::regex const re(o.filter);
    std::vector<Result> results;
    std::cout << std::left << std::setw(28) << "Benchmark" << std::right << std::setw(16) << "Time"
              << std::setw(16) << "CPU" << std::setw(14) << "Iterations" << "  UserCounters...\n";
    for (Case const& c : cases<T, IndexType>(ex)) {
        if (!std::regex_search(c.name, re)) {
            continue;
        }
        Result const r = run_case(c, o.min_time);
        results.push_back(r);
        std::ostringstream line;
        line << std::left << std::setw(28) << r.name << std::right << std::fixed << std::setprecision(0)
             << std::setw(13) << r.real_ns << " ns" << std::setw(13) << r.cpu_ns << " ns" << std::setw(14)
             << r.iterations;
        line << std::setprecision(4) << std::scientific;
        if (r.bytes > 0) {
            line << "  bytes_per_second=" << r.bytes / (r.cpu_ns * 1e-9);
        }
        if (r.cells > 0) {
            line << "  cells_per_second=" << r.cells / (r.cpu_ns * 1e-9);
        }
        std::cout << line.str() << '\n' << std::flush;
    }
    if (!o.out.empty()) {
        std::ofstream f(o.out);
        if (!f) {
            throw std::runtime_error("cannot write " + o.out);
        }
        write_json(f, results, o, exe);
    }
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    try {
        Options const o = parse(argc, argv);
        // spin while waiting on the GPU, so that CPU time measures the iteration
        if (hipSetDeviceFlags(hipDeviceScheduleSpin) != hipSuccess) {
            throw std::runtime_error("hipSetDeviceFlags failed");
        }
        if (o.real == "f64" && o.index == "int") {
            return run<double, int>(o, argv[0]);
        }
        if (o.real == "f64" && o.index == "long") {
            return run<double, std::int64_t>(o, argv[0]);
        }
        if (o.real == "f32" && o.index == "int") {
            return run<float, int>(o, argv[0]);
        }
        if (o.real == "f32" && o.index == "long") {
            return run<float, std::int64_t>(o, argv[0]);
        }
        throw std::invalid_argument("--real=f64|f32, --index=int|long");
    } catch (std::exception const& e) {
        std::cerr << "euler_benchmarks: " << e.what() << '\n';
        return 1;
    }
}
