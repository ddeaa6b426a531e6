// logging.cpp -- session state and the rocsolver_log_* C API of the
// multi-level logging (see logging.hpp).  Reference behaviour followed:
// library/src/common/rocsolver_logger.cpp:191-231 (log_begin: ROCSOLVER_LAYER
// and ROCSOLVER_LEVELS parsed with strtol, output streams from
// ROCSOLVER_LOG_{TRACE,BENCH,PROFILE}_PATH or ROCSOLVER_LOG_PATH, else
// stderr, a version header once per file), :232-330 (end, set_layer_mode,
// set_max_levels >= 1, restore_defaults, write/flush_profile; every call
// without an active session returns rocblas_status_internal_error), and
// library/src/include/rocsolver_logger.hpp:160-300 (bench line, trace tree,
// profile tree of the nested functions only, timed after synchronising the
// handle's stream).
#include "logging.hpp"

#include "../../include/rocsolver_amd.h"

#include <hip/hip_runtime.h>

#include <cerrno>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>


namespace rsa
{
namespace logging
{

std::atomic<uint32_t> g_mode{0};

void Args::add(const char* flag, long long v)
{
    bench += " ";
    bench += flag;
    bench += " " + std::to_string(v);
    std::string key = flag;
    while(!key.empty() && key[0] == '-')
        key.erase(0, 1);
    if(!trace.empty())
        trace += ", ";
    trace += key + ": " + std::to_string(v);
}

void Args::add(const char* flag, char v)
{
    bench += " ";
    bench += flag;
    bench += " ";
    bench += v;
    std::string key = flag;
    while(!key.empty() && key[0] == '-')
        key.erase(0, 1);
    if(!trace.empty())
        trace += ", ";
    trace += key + ": " + std::string(1, v);
}

namespace
{
struct Entry
{
    std::string name;
    double start_us;
    std::vector<std::string> callers; // names of levels 1 .. level-1
};

struct ProfNode
{
    long calls = 0;
    double time_us = 0;
    std::map<std::string, ProfNode> nested;
};

struct Session
{
    uint32_t mode = 0;
    int max_levels = 1;
    FILE* trace = stderr;
    FILE* bench = stderr;
    FILE* profile = stderr;
    std::vector<std::pair<std::string, FILE*>> files;
    std::map<rocblas_handle, std::vector<Entry>> stacks;
    std::map<std::string, ProfNode> prof;
    std::string trace_buf;
};

std::mutex g_mutex;
Session* g_session = nullptr;

double now_us()
{
    using namespace std::chrono;
    return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

FILE* open_stream(Session& s, const char* var)
{
    const char* path = std::getenv(var);
    if(!path)
        path = std::getenv("ROCSOLVER_LOG_PATH");
    if(!path)
        return stderr;
    for(auto& f : s.files)
        if(f.first == path)
            return f.second;
    FILE* fp = std::fopen(path, "w");
    if(!fp)
        return nullptr;
    char ver[64] = "unknown";
    rocsolver_get_version_string(ver, sizeof(ver));
    size_t rbs = 0;
    std::string rb = "unknown";
    if(rocblas_get_version_string_size(&rbs) == rocblas_status_success && rbs > 1)
    {
        rb.assign(rbs, '\0');
        rocblas_get_version_string(&rb[0], rbs);
        rb.resize(rbs - 1);
    }
    std::fprintf(fp, "ROCSOLVER LOG FILE\nrocSOLVER Version: %s\nrocBLAS Version: %s\n", ver,
                 rb.c_str());
    std::fflush(fp);
    s.files.emplace_back(path, fp);
    return fp;
}

void append_profile(std::string& out, const std::map<std::string, ProfNode>& m, int level,
                    int max_levels)
{
    for(const auto& kv : m)
    {
        const ProfNode& e = kv.second;
        char buf[512];
        std::snprintf(buf, sizeof(buf), "%*s%s: Calls: %ld, Total Time: %.3f ms", 4 * (level - 1),
                      "", kv.first.c_str(), e.calls, e.time_us * 1e-3);
        out += buf;
        if(!e.nested.empty())
        {
            double t = 0;
            for(const auto& n : e.nested)
                t += n.second.time_us;
            std::snprintf(buf, sizeof(buf), " (in nested functions: %.3f ms)\n", t * 1e-3);
            out += buf;
            if(level < max_levels)
                append_profile(out, e.nested, level + 1, max_levels);
        }
        else
            out += '\n';
    }
}

void print_profile(Session& s, bool clear)
{
    if((s.mode & PROFILE) && !s.prof.empty())
    {
        std::string str;
        append_profile(str, s.prof, 1, s.max_levels);
        std::fprintf(s.profile, "------- PROFILE -------\n%s\n", str.c_str());
        std::fflush(s.profile);
        if(clear)
            s.prof.clear();
    }
}
} // namespace

Scope::Scope(rocblas_handle handle, bool top, const std::string& name, char prec,
             const Args& args)
    : handle_(handle)
    , top_(top)
    , live_(false)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    Session* s = g_session;
    if(!s || !(s->mode & (TRACE | BENCH | PROFILE)))
        return;
    live_ = true;
    std::vector<Entry>& stack = s->stacks[handle];
    Entry e;
    e.name = top ? std::string("rocsolver_") + prec + name : name;
    e.start_us = now_us();
    for(size_t i = 1; i < stack.size(); ++i)
        e.callers.push_back(stack[i].name);
    const int level = int(stack.size());
    if(top)
    {
        if(s->mode & BENCH)
        {
            std::fprintf(s->bench, "./rocsolver-bench -f %s -r %c%s\n", name.c_str(), prec,
                         args.bench.c_str());
            std::fflush(s->bench);
        }
        if(s->mode & TRACE)
            s->trace_buf += "------- ENTER " + e.name + " trace tree -------\n";
    }
    else if((s->mode & TRACE) && level <= s->max_levels)
    {
        s->trace_buf += std::string(4 * (level - 1), ' ') + name;
        if(!args.trace.empty())
            s->trace_buf += " (" + args.trace + ")";
        s->trace_buf += '\n';
    }
    stack.push_back(std::move(e));
}

Scope::~Scope()
{
    if(!live_)
        return;
    bool profile = false;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        profile = g_session && (g_session->mode & PROFILE) && !top_;
    }
    double end = 0;
    if(profile)
    {
        hipStream_t stream = nullptr;
        rocblas_get_stream(handle_, &stream);
        (void)hipStreamSynchronize(stream);
        end = now_us();
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    Session* s = g_session;
    if(!s)
        return;
    auto it = s->stacks.find(handle_);
    if(it == s->stacks.end() || it->second.empty())
        return;
    Entry e = std::move(it->second.back());
    it->second.pop_back();
    if(it->second.empty())
        s->stacks.erase(it);
    if(profile)
    {
        std::map<std::string, ProfNode>* m = &s->prof;
        for(const std::string& c : e.callers)
            m = &(*m)[c].nested;
        ProfNode& node = (*m)[e.name];
        node.calls++;
        node.time_us += end - e.start_us;
    }
    if(top_ && (s->mode & TRACE))
    {
        s->trace_buf += "------- EXIT " + e.name + " trace tree -------\n\n";
        std::fputs(s->trace_buf.c_str(), s->trace);
        std::fflush(s->trace);
        s->trace_buf.clear();
    }
}

} // namespace logging
} // namespace rsa

using namespace rsa::logging;

extern "C" {

rocblas_status rocsolver_log_begin(void)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if(g_session)
        return rocblas_status_internal_error;
    Session* s = g_session = new Session();
    if(const char* str = std::getenv("ROCSOLVER_LAYER"))
    {
        errno = 0;
        const long v = std::strtol(str, nullptr, 0);
        if(errno || v < 0 || uint64_t(v) > uint64_t(UINT32_MAX))
            return rocblas_status_internal_error;
        s->mode = uint32_t(v);
    }
    if(const char* str = std::getenv("ROCSOLVER_LEVELS"))
    {
        errno = 0;
        const long v = std::strtol(str, nullptr, 0);
        if(errno || v < 1 || v > long(INT_MAX))
            return rocblas_status_internal_error;
        s->max_levels = int(v);
    }
    s->trace = open_stream(*s, "ROCSOLVER_LOG_TRACE_PATH");
    s->bench = open_stream(*s, "ROCSOLVER_LOG_BENCH_PATH");
    s->profile = open_stream(*s, "ROCSOLVER_LOG_PROFILE_PATH");
    if(!s->trace || !s->bench || !s->profile)
    {
        s->mode = 0;
        return rocblas_status_internal_error;
    }
    g_mode.store(s->mode);
    return rocblas_status_success;
}

rocblas_status rocsolver_log_end(void)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    Session* s = g_session;
    if(!s || !s->stacks.empty())
        return rocblas_status_internal_error;
    print_profile(*s, false);
    for(auto& f : s->files)
        std::fclose(f.second);
    delete s;
    g_session = nullptr;
    g_mode.store(0);
    return rocblas_status_success;
}

rocblas_status rocsolver_log_set_layer_mode(const rocblas_layer_mode_flags layer_mode)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if(!g_session)
        return rocblas_status_internal_error;
    g_session->mode = uint32_t(layer_mode);
    g_mode.store(g_session->mode);
    return rocblas_status_success;
}

rocblas_status rocsolver_log_set_max_levels(const rocblas_int max_levels)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if(!g_session)
        return rocblas_status_internal_error;
    if(max_levels < 1)
        return rocblas_status_invalid_value;
    g_session->max_levels = max_levels;
    return rocblas_status_success;
}

rocblas_status rocsolver_log_restore_defaults(void)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if(!g_session)
        return rocblas_status_internal_error;
    g_session->max_levels = 1;
    g_session->mode = 0;
    g_mode.store(0);
    return rocblas_status_success;
}

rocblas_status rocsolver_log_write_profile(void)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if(!g_session)
        return rocblas_status_internal_error;
    print_profile(*g_session, false);
    return rocblas_status_success;
}

rocblas_status rocsolver_log_flush_profile(void)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if(!g_session)
        return rocblas_status_internal_error;
    print_profile(*g_session, true);
    return rocblas_status_success;
}

} // extern "C"
