// C-ABI plumbing: thread-local error messages, version, and the step-body
// graph capture helpers (hipStreamBeginCapture / hipGraphInstantiate).
#include <cstdarg>
#include <cstdio>

#include "common.h"

namespace glide {

static thread_local char g_err[1024] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory) ? GLIDE_ERR_MEMORY : GLIDE_ERR_DEVICE;
    return set_error(code, "%s failed: %s (%d)", what, hipGetErrorString(e), static_cast<int>(e));
}

}  // namespace glide

using namespace glide;

extern "C" int glide_abi_version(void) { return GLIDE_ABI_VERSION; }

extern "C" const char* glide_last_error(void) { return g_err; }

extern "C" int glide_graph_begin(void* stream) {
    GLIDE_REQUIRE(stream != nullptr, "graph capture needs a non-default stream");
    GLIDE_HIP(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
    return GLIDE_OK;
}

struct GraphHandle {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" int glide_graph_end(void* stream, void** handle) {
    GLIDE_REQUIRE(stream != nullptr && handle != nullptr, "NULL stream/handle");
    hipGraph_t g = nullptr;
    GLIDE_HIP(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return check_hip(e, "hipGraphInstantiate");
    }
    auto* h = new GraphHandle{g, ex};
    *handle = h;
    return GLIDE_OK;
}

extern "C" int glide_graph_launch(void* handle, void* stream) {
    GLIDE_REQUIRE(handle != nullptr, "NULL graph handle");
    auto* h = static_cast<GraphHandle*>(handle);
    GLIDE_HIP(hipGraphLaunch(h->exec, static_cast<hipStream_t>(stream)));
    return GLIDE_OK;
}

extern "C" int glide_graph_destroy(void* handle) {
    if (handle == nullptr) return GLIDE_OK;
    auto* h = static_cast<GraphHandle*>(handle);
    (void)hipGraphExecDestroy(h->exec);
    (void)hipGraphDestroy(h->graph);
    delete h;
    return GLIDE_OK;
}
