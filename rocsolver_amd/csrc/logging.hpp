// logging.hpp -- the multi-level logging of the public entry points (trace,
// bench and profile output, selected by rocblas_layer_mode flags), with the
// observable behaviour of the reference's logger
// (library/src/common/rocsolver_logger.cpp, library/src/include/
// rocsolver_logger.hpp): the rocsolver_log_* API, the ROCSOLVER_LAYER /
// ROCSOLVER_LEVELS / ROCSOLVER_LOG_*_PATH environment variables, bench lines
// "./rocsolver-bench -f <function> -r <precision> <arguments>", trace trees
// and per-function profile totals.  Nothing is logged (and nothing costs
// more than one atomic load) unless a session is active with a mode set.
#pragma once

#include <rocblas/rocblas.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <type_traits>

namespace rsa
{
namespace logging
{

// layer-mode bits (rocblas_layer_mode of rocblas-types.h)
constexpr uint32_t TRACE = 0x1;
constexpr uint32_t BENCH = 0x2;
constexpr uint32_t PROFILE = 0x4;

extern std::atomic<uint32_t> g_mode; // current mode, 0 without an active session

inline bool enabled()
{
    return g_mode.load(std::memory_order_relaxed) & (TRACE | BENCH | PROFILE);
}

// "name value" pairs of a call, in order
struct Args
{
    std::string bench; // " -m 32 -n 32 ..."
    std::string trace; // "m: 32, n: 32, ..."
    void add(const char* flag, long long v);
    void add(const char* flag, char v);
};

inline char op_char(rocblas_operation op)
{
    return op == rocblas_operation_transpose ? 'T'
                                             : (op == rocblas_operation_conjugate_transpose ? 'C' : 'N');
}
inline char fill_char(rocblas_fill f)
{
    return f == rocblas_fill_upper ? 'U' : (f == rocblas_fill_lower ? 'L' : 'F');
}

inline void collect(Args&) {}
template <typename V, typename... Rest>
void collect(Args& a, const char* flag, V v, Rest... rest)
{
    if constexpr(std::is_same<V, rocblas_operation>::value)
        a.add(flag, op_char(v));
    else if constexpr(std::is_same<V, rocblas_fill>::value)
        a.add(flag, fill_char(v));
    else
        a.add(flag, (long long)v);
    collect(a, rest...);
}

// Entry of a public function (top level: bench line, trace tree) or of an
// internal stage (nested: trace line, profile entry).  The destructor ends
// it (trace tree closed, profile time taken after synchronising the
// handle's stream, as the reference's profile mode does).
class Scope
{
public:
    Scope(rocblas_handle handle, bool top, const std::string& name, char prec, const Args& args);
    ~Scope();
    Scope(const Scope&) = delete;
    Scope& operator=(const Scope&) = delete;

private:
    rocblas_handle handle_;
    bool top_;
    bool live_;
};

} // namespace logging
} // namespace rsa

// Logs a public entry point for the rest of the enclosing scope: P is the
// precision letter (s, d), NAME the function name without "rocsolver_<P>",
// then "flag", value pairs as the reference's bench lines spell them.
#define RSA_LOG_TOP(HANDLE, P, NAME, ...)                                                      \
    std::unique_ptr<rsa::logging::Scope> rsa_log_scope_;                                       \
    if(rsa::logging::enabled())                                                               \
    {                                                                                          \
        rsa::logging::Args rsa_log_args_;                                                      \
        rsa::logging::collect(rsa_log_args_, __VA_ARGS__);                                     \
        rsa_log_scope_ = std::make_unique<rsa::logging::Scope>(HANDLE, true, NAME, #P[0],      \
                                                               rsa_log_args_);                 \
    }

// Logs an internal stage (the kernel driver chosen) for the rest of the scope.
#define RSA_LOG_STAGE(HANDLE, NAME, ...)                                                       \
    std::unique_ptr<rsa::logging::Scope> rsa_log_stage_;                                       \
    if(rsa::logging::enabled())                                                               \
    {                                                                                          \
        rsa::logging::Args rsa_log_args_;                                                      \
        rsa::logging::collect(rsa_log_args_, __VA_ARGS__);                                     \
        rsa_log_stage_ = std::make_unique<rsa::logging::Scope>(HANDLE, false, NAME, 0,         \
                                                               rsa_log_args_);                 \
    }
