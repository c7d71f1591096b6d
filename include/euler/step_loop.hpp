// Host-sync-free step loop of the reference driver (simulations/euler_simulation.cpp:50-67):
// one fused launch per step (godunov + cons_to_prim + the next time step, dt kept
// on the device) over double-buffered primitives, optionally replayed from HIP
// graphs of 2*kChunkPairs steps and of two steps. Semantics as in the reference
// loop: the first dt
// comes from the initial primitives (ghosts included), step 0 reads the initial
// ghost primitives and later steps read the ghost primitives round-tripped
// through cons_to_prim; simulated time is accumulated in FP64 on the device.
#pragma once

#include <utility>

#include <hip/hip_runtime_api.h>

#include "detail_capi.hpp"
#include "perfect_gas.hpp"
#include "time_step.hpp"
#include "uniform_mesh.hpp"

namespace eb {

template <class T, class IndexType>
class StepLoop
{
public:
    using prim_view = EulerPrimArrays<mdspan<T, dextents<IndexType, 3>, layout_left>>;
    using cons_view = EulerConsArrays<mdspan<T, dextents<IndexType, 3>, layout_left>>;

    // prim holds the initial state (and, after finish(), the current one); cons
    // must already hold prim_to_cons(prim).
    StepLoop(HipExecSpace const& exec_space, prim_view const& prim, cons_view const& cons, PerfectGas<T> const& eos,
             UniformMesh3d<T> const& mesh, T cfl)
        : m_ex(exec_space), m_cons(cons), m_eos(eos), m_mesh(mesh), m_cfl(cfl)
    {
        detail::check_pair(prim, cons);
        for (std::size_t r = 0; r < 3; ++r) {
            if (prim.d.extent(r) < 3) {
                throw Error(EB_ERR_INVALID_ARGUMENT, "the step loop needs at least one interior cell");
            }
        }
        std::size_t const n = size(prim);
        m_alloc = create_prim_arrays_1d<T>(exec_space, n);
        m_buf[0] = prim;
        m_buf[1] = to_mdspan<mdspan<T, dextents<IndexType, 3>, layout_left>>(m_alloc, prim.d.extent(0),
                                                                              prim.d.extent(1), prim.d.extent(2));
        m_dt = View<T>(exec_space, "dt", 3); // [0], [1]: ping-pong raw dt; [2]: ghost-cell minimum
        m_time = View<double>(exec_space, "time", 1);
    }

    StepLoop(StepLoop const&) = delete;
    StepLoop& operator=(StepLoop const&) = delete;

    ~StepLoop()
    {
        for (int c = 0; c < 2; ++c) {
            if (m_graph_exec[c] != nullptr) {
                (void)hipGraphExecDestroy(m_graph_exec[c]);
            }
            if (m_graph[c] != nullptr) {
                (void)hipGraphDestroy(m_graph[c]);
            }
        }
    }

    // Steps per replay of the long graph: short graphs (two steps) measured ~4%
    // slower than eager launches at 128^3, so bulk steps replay a longer one.
    static constexpr int kChunkPairs = 16;

    void start()
    {
        time_step(m_ex, as_const(m_buf[0]), m_eos, m_mesh, m_dt.data());
        ghost_refresh(1);
        m_started = true;
    }

    void step()
    {
        if (!m_started) {
            start();
        }
        int const src = static_cast<int>(m_steps % 2);
        fused(src, m_ex.stream());
        if (m_steps == 0) {
            ghost_refresh(0);
        }
        ++m_steps;
    }

    // Advance nsteps; with use_graph, the steps after the first two are replayed
    // from captured HIP graphs (runs of 2*kChunkPairs steps, then pairs).
    void run(long nsteps, bool use_graph = false)
    {
        while (m_steps < 2 && nsteps > 0) {
            step();
            --nsteps;
        }
        if (use_graph && nsteps >= 2 && m_steps % 2 == 1) {
            step(); // the captured pairs start from buffer 0
            --nsteps;
        }
        if (use_graph && nsteps >= 2 * kChunkPairs) {
            capture(1, kChunkPairs);
            while (nsteps >= 2 * kChunkPairs) {
                detail::check_hip(hipGraphLaunch(m_graph_exec[1], m_ex.stream()), "hipGraphLaunch");
                m_steps += 2 * kChunkPairs;
                nsteps -= 2 * kChunkPairs;
            }
        }
        if (use_graph && nsteps >= 2) {
            capture(0, 1);
            while (nsteps >= 2) {
                detail::check_hip(hipGraphLaunch(m_graph_exec[0], m_ex.stream()), "hipGraphLaunch");
                m_steps += 2;
                nsteps -= 2;
            }
        }
        while (nsteps > 0) {
            step();
            --nsteps;
        }
    }

    long steps_done() const noexcept
    {
        return m_steps;
    }

    prim_view const& current_prim() const noexcept
    {
        return m_buf[m_steps % 2];
    }

    // Simulated time sum(cfl * dt) (blocking).
    double time() const
    {
        double t = 0;
        detail::check_hip(hipMemcpyAsync(&t, m_time.data(), sizeof(double), hipMemcpyDeviceToHost, m_ex.stream()),
                          "hipMemcpyAsync");
        m_ex.fence();
        return t;
    }

    // Raw dt the next step will use (blocking).
    T next_dt() const
    {
        T dt {};
        detail::check_hip(hipMemcpyAsync(&dt, m_dt.data() + m_steps % 2, sizeof(T), hipMemcpyDeviceToHost,
                                         m_ex.stream()),
                          "hipMemcpyAsync");
        m_ex.fence();
        return dt;
    }

    // Leave the current primitives in the caller's prim arrays.
    void finish()
    {
        if (m_steps % 2 == 1) {
            auto const dst = detail::ptrs(m_buf[0]);
            auto const src = detail::ptrs(m_buf[1]);
            std::size_t const bytes = size(m_buf[0]) * sizeof(T);
            T* const d[5] = {dst.d, dst.p, dst.ux0, dst.ux1, dst.ux2};
            T* const s[5] = {src.d, src.p, src.ux0, src.ux1, src.ux2};
            for (int f = 0; f < 5; ++f) {
                detail::check_hip(hipMemcpyAsync(d[f], s[f], bytes, hipMemcpyDeviceToDevice, m_ex.stream()),
                                  "hipMemcpyAsync");
            }
        }
        m_ex.fence();
    }

private:
    void ghost_refresh(int target)
    {
        auto const c = detail::ptrs(m_cons);
        auto const p = detail::ptrs(m_buf[target]);
        detail::check(detail::capi<T>::ghost_refresh(m_ex.context(), m_ex.stream(), &c, &p, detail::ext(m_cons, 0),
                                                     detail::ext(m_cons, 1), detail::ext(m_cons, 2),
                                                     detail::index_bits<IndexType>(m_cons), m_eos.gamma(), m_mesh.dx0(),
                                                     m_mesh.dx1(), m_mesh.dx2(), m_dt.data() + 2));
    }

    void fused(int src, hipStream_t stream)
    {
        int const dst = 1 - src;
        auto const pin = detail::ptrs(m_buf[src]);
        auto const c = detail::ptrs(m_cons);
        auto const pout = detail::ptrs(m_buf[dst]);
        detail::check(detail::capi<T>::step_fused(
                m_ex.context(), stream, &pin, &c, &pout, detail::ext(m_cons, 0), detail::ext(m_cons, 1),
                detail::ext(m_cons, 2), detail::index_bits<IndexType>(m_cons), m_eos.gamma(), m_mesh.dx0(), m_mesh.dx1(),
                m_mesh.dx2(), m_cfl, m_dt.data() + src, m_dt.data() + 2, m_dt.data() + dst, m_time.data()));
    }

    // Graph c (0: one pair of steps, 1: kChunkPairs pairs), captured on first use.
    void capture(int c, int pairs)
    {
        if (m_graph_exec[c] != nullptr) {
            return;
        }
        hipStream_t const s = m_ex.stream();
        detail::check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            for (int p = 0; p < pairs; ++p) {
                fused(0, s);
                fused(1, s);
            }
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g != nullptr) {
                (void)hipGraphDestroy(g);
            }
            throw;
        }
        detail::check_hip(hipStreamEndCapture(s, &m_graph[c]), "hipStreamEndCapture");
        detail::check_hip(hipGraphInstantiate(&m_graph_exec[c], m_graph[c], nullptr, nullptr, 0),
                          "hipGraphInstantiate");
    }

    HipExecSpace m_ex;
    cons_view m_cons;
    PerfectGas<T> m_eos;
    UniformMesh3d<T> m_mesh;
    T m_cfl;
    EulerPrimArrays<View<T>> m_alloc;
    prim_view m_buf[2];
    View<T> m_dt;
    View<double> m_time;
    long m_steps = 0;
    bool m_started = false;
    hipGraph_t m_graph[2] = {nullptr, nullptr};
    hipGraphExec_t m_graph_exec[2] = {nullptr, nullptr};
};

} // namespace eb
