// Execution space of the C++ API: a HIP stream plus the eb_context that holds the
// device scratch of the single-pass CFL reduction. It takes the place of
// Kokkos::DefaultExecutionSpace in the reference's operator signatures
// (euler_operators/godunov.hpp:16). Copies share one stream and context.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime_api.h>

#include "../euler_capi.h"

namespace eb {

class Error : public std::runtime_error
{
public:
    Error(int status, std::string const& what) : std::runtime_error(what), m_status(status) {}

    int status() const noexcept
    {
        return m_status;
    }

private:
    int m_status;
};

namespace detail {

inline void check(int status)
{
    if (status != EB_OK) {
        throw Error(status, std::string(eb_status_string(status)) + ": " + eb_last_error());
    }
}

inline void check_hip(hipError_t e, char const* what)
{
    if (e != hipSuccess) {
        throw Error(EB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

} // namespace detail

class HipExecSpace
{
    struct State
    {
        hipStream_t stream = nullptr;
        eb_context* ctx = nullptr;
        bool owns_stream = false;

        ~State()
        {
            if (ctx != nullptr) {
                (void)eb_context_destroy(ctx);
            }
            if (owns_stream && stream != nullptr) {
                (void)hipStreamDestroy(stream);
            }
        }
    };

public:
    // A new non-blocking stream (default priority) on the current device.
    HipExecSpace() : m_state(std::make_shared<State>())
    {
        detail::check_hip(hipStreamCreateWithFlags(&m_state->stream, hipStreamNonBlocking), "hipStreamCreate");
        m_state->owns_stream = true;
        detail::check(eb_context_create(&m_state->ctx));
    }

    // Work is enqueued on a stream the caller owns (e.g. one from torch).
    explicit HipExecSpace(hipStream_t stream) : m_state(std::make_shared<State>())
    {
        m_state->stream = stream;
        detail::check(eb_context_create(&m_state->ctx));
    }

    hipStream_t stream() const noexcept
    {
        return m_state->stream;
    }

    eb_context* context() const noexcept
    {
        return m_state->ctx;
    }

    void fence() const
    {
        detail::check_hip(hipStreamSynchronize(m_state->stream), "hipStreamSynchronize");
    }

private:
    std::shared_ptr<State> m_state;
};

using DefaultExecutionSpace = HipExecSpace;

} // namespace eb
