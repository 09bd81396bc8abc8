// Handle lifecycle, stream binding, workspace policy, version and logging.
//
// Behaviour follows reference blas/common/helper/aclblas_auxiliary.cpp:24-148
// and aclblas_handle_internal.h:95-209; the logger entry points declared at
// reference include/cann_ops_blas.h:116-133 are defined here with C linkage.
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "internal.h"

namespace aclblas {
namespace {

struct LoggerState {
    std::mutex mu;
    bool on = false;
    bool to_stdout = false;
    bool to_stderr = true;
    const char* file = nullptr;
    aclblasLogCallback callback = nullptr;
    int level = 0;  // 0 = errors only when on, 1 = info too
    bool env_read = false;
};

LoggerState& Logger()
{
    static LoggerState s;
    return s;
}

void ReadEnvOnce(LoggerState& s)
{
    if (s.env_read) {
        return;
    }
    s.env_read = true;
    const char* lvl = std::getenv("ACLBLAS_LOG_LEVEL");
    if (lvl != nullptr && lvl[0] != '\0') {
        int v = std::atoi(lvl);
        if (v > 0) {
            s.on = true;
            s.level = v > 1 ? 1 : 0;
        }
    }
}

void Emit(int level, const char* api, const char* fmt, va_list ap)
{
    LoggerState& s = Logger();
    std::lock_guard<std::mutex> lock(s.mu);
    ReadEnvOnce(s);
    if (!s.on && s.callback == nullptr) {
        return;
    }
    if (level > s.level && s.callback == nullptr) {
        return;
    }
    char body[512];
    std::vsnprintf(body, sizeof(body), fmt, ap);
    char line[640];
    std::snprintf(line, sizeof(line), "[aclblas][%s] %s: %s", level == 0 ? "E" : "I", api, body);
    if (s.callback != nullptr) {
        s.callback(line);
    }
    if (!s.on) {
        return;
    }
    if (s.to_stdout) {
        std::fprintf(stdout, "%s\n", line);
    }
    if (s.to_stderr) {
        std::fprintf(stderr, "%s\n", line);
    }
    if (s.file != nullptr) {
        FILE* f = std::fopen(s.file, "a");
        if (f != nullptr) {
            std::fprintf(f, "%s\n", line);
            std::fclose(f);
        }
    }
}

// The handle's own device arena (32 MiB at creation, grown on demand).
aclblasStatus_t AcquireOwnedArena(_aclblas_handle* h, size_t size)
{
    if (size == 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    void* p = nullptr;
    if (hipMalloc(&p, size) != hipSuccess) {
        (void)hipGetLastError();
        return ACLBLAS_STATUS_ALLOC_FAILED;
    }
    h->workspace = p;
    h->workspace_size = size;
    h->owned_size_hint = size;
    h->owner = WorkspaceOwner::Library;
    h->arena_in_graph = false;
    return ACLBLAS_STATUS_SUCCESS;
}

// Gives up the handle's own arena.  An arena that a graph captured on this
// handle may still address is parked until aclblasDestroy instead of freed, so
// replaying such a graph after the arena was replaced never touches freed
// memory.
void ReleaseOwnedArena(_aclblas_handle* h)
{
    if (h->owner != WorkspaceOwner::Library || h->workspace == nullptr) {
        return;
    }
    if (h->arena_in_graph) {
        h->parked_arenas.push_back(h->workspace);
    } else {
        (void)hipFree(h->workspace);
    }
    h->workspace = nullptr;
    h->workspace_size = 0;
    h->arena_in_graph = false;
}

// After a borrowed workspace or a stream change the handle works in an arena
// of its own again, of the size it last had.
aclblasStatus_t ReturnToOwnedArena(_aclblas_handle* h)
{
    if (!HandleHasDevice(h)) {
        h->workspace = nullptr;
        h->workspace_size = 0;
        h->owner = WorkspaceOwner::Library;
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (h->owner == WorkspaceOwner::Library && h->workspace != nullptr) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    aclblasStatus_t st = SyncHandleStream(h);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    h->workspace = nullptr;
    h->workspace_size = 0;
    h->owner = WorkspaceOwner::Library;
    size_t size = h->owned_size_hint > 0 ? h->owned_size_hint : kDefaultWorkspaceSize;
    return AcquireOwnedArena(h, size);
}

}  // namespace

void LogError(const char* api, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    Emit(0, api, fmt, ap);
    va_end(ap);
}

void LogInfo(const char* api, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    Emit(1, api, fmt, ap);
    va_end(ap);
}

aclblasStatus_t SyncHandleStream(_aclblas_handle* h)
{
    if (h == nullptr || !HandleHasDevice(h)) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    return hipStreamSynchronize(h->stream) == hipSuccess ? ACLBLAS_STATUS_SUCCESS
                                                         : ACLBLAS_STATUS_EXECUTION_FAILED;
}

aclblasStatus_t EnsureWorkspace(_aclblas_handle* h, size_t bytes)
{
    if (h == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (!HandleHasDevice(h)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    const bool capturing = StreamCapturing(h);
    if (h->owner == WorkspaceOwner::User) {
        if (bytes > h->workspace_size) {
            LogError("aclblasHandle", "the caller's workspace holds %zu bytes, this call needs %zu",
                     h->workspace_size, bytes);
            return ACLBLAS_STATUS_ALLOC_FAILED;
        }
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (bytes <= h->workspace_size) {
        if (capturing) {
            h->arena_in_graph = true;
        }
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (bytes > kMaxWorkspaceSize) {
        LogError("aclblasHandle", "workspace of %zu bytes exceeds the %zu byte limit", bytes, kMaxWorkspaceSize);
        return ACLBLAS_STATUS_ALLOC_FAILED;
    }
    if (capturing) {
        LogError("aclblasHandle", "workspace must grow to %zu bytes, which cannot happen during stream capture: "
                 "run the call once before capturing or set a workspace", bytes);
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    aclblasStatus_t st = SyncHandleStream(h);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    size_t doubled = h->workspace_size > 0 ? h->workspace_size * 2 : kDefaultWorkspaceSize;
    size_t size = bytes > doubled ? bytes : doubled;
    if (size > kMaxWorkspaceSize) {
        size = kMaxWorkspaceSize;
    }
    ReleaseOwnedArena(h);
    return AcquireOwnedArena(h, size);
}

void MarkWorkspaceUse(_aclblas_handle* h)
{
    if (h->owner == WorkspaceOwner::Library && h->workspace != nullptr && StreamCapturing(h)) {
        h->arena_in_graph = true;
    }
}

bool StreamCapturing(const _aclblas_handle* h)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) != hipSuccess) {
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

aclblasStatus_t StageToWorkspace(_aclblas_handle* h, const void* src, size_t bytes, void** dev)
{
    aclblasStatus_t st = EnsureWorkspace(h, bytes);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    if (StreamCapturing(h)) {
        // A captured copy re-reads its host source at every replay: give it a
        // pinned buffer of its own that lives as long as the handle.
        void* keep = nullptr;
        if (hipHostMalloc(&keep, bytes, hipHostMallocDefault) != hipSuccess) {
            LogError("aclblasHandle", "pinned staging allocation of %zu bytes failed", bytes);
            return ACLBLAS_STATUS_ALLOC_FAILED;
        }
        h->captured_stages.push_back(keep);
        std::memcpy(keep, src, bytes);
        if (hipMemcpyAsync(h->workspace, keep, bytes, hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            return ACLBLAS_STATUS_EXECUTION_FAILED;
        }
        *dev = h->workspace;
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (h->stage_event == nullptr) {
        if (hipEventCreateWithFlags(&h->stage_event, hipEventDisableTiming) != hipSuccess) {
            h->stage_event = nullptr;
            return ACLBLAS_STATUS_EXECUTION_FAILED;
        }
    } else if (hipEventSynchronize(h->stage_event) != hipSuccess) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    if (bytes > h->host_stage_size) {
        if (h->host_stage != nullptr) {
            (void)hipHostFree(h->host_stage);
            h->host_stage = nullptr;
            h->host_stage_size = 0;
        }
        size_t size = bytes < 65536 ? 65536 : bytes;
        if (hipHostMalloc(&h->host_stage, size, hipHostMallocDefault) != hipSuccess) {
            h->host_stage = nullptr;
            LogError("aclblasHandle", "pinned staging allocation of %zu bytes failed", size);
            return ACLBLAS_STATUS_ALLOC_FAILED;
        }
        h->host_stage_size = size;
    }
    std::memcpy(h->host_stage, src, bytes);
    if (hipMemcpyAsync(h->workspace, h->host_stage, bytes, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipEventRecord(h->stage_event, h->stream) != hipSuccess) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    *dev = h->workspace;
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace aclblas

using namespace aclblas;

extern "C" {

aclblasStatus_t aclblasCreate(aclblasHandle_t* handle)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (*handle != nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    auto* h = new (std::nothrow) _aclblas_handle();
    if (h == nullptr) {
        return ACLBLAS_STATUS_ALLOC_FAILED;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        // No GPU visible: a handle without workspace, usable for argument checks.
        *handle = h;
        return ACLBLAS_STATUS_SUCCESS;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete h;
        return ACLBLAS_STATUS_NOT_INITIALIZED;
    }
    h->device = dev;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        cus = 256;
    }
    h->cu_count = cus;
    aclblasStatus_t st = AcquireOwnedArena(h, kDefaultWorkspaceSize);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        delete h;
        return st;
    }
    *handle = h;
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasDestroy(aclblasHandle_t handle)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (SyncHandleStream(handle) != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasDestroy", "stream synchronisation failed before handle destruction");
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    handle->arena_in_graph = false;
    ReleaseOwnedArena(handle);
    for (void* p : handle->parked_arenas) {
        (void)hipFree(p);
    }
    if (handle->host_stage != nullptr) {
        (void)hipHostFree(handle->host_stage);
    }
    if (handle->stage_event != nullptr) {
        (void)hipEventDestroy(handle->stage_event);
    }
    for (void* p : handle->captured_stages) {
        (void)hipHostFree(p);
    }
    delete handle;
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasSetStream(aclblasHandle_t handle, aclrtStream stream)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    aclblasStatus_t st = SyncHandleStream(handle);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    handle->stream = reinterpret_cast<hipStream_t>(stream);
    return ReturnToOwnedArena(handle);
}

aclblasStatus_t aclblasGetStream(aclblasHandle_t handle, aclrtStream* stream)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (stream == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    *stream = reinterpret_cast<aclrtStream>(handle->stream);
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasSetWorkspace(aclblasHandle_t handle, void* workspace, size_t workspaceSize)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (workspace == nullptr || workspaceSize == 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (handle->owner == WorkspaceOwner::User && workspaceSize <= handle->workspace_size) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    aclblasStatus_t st = SyncHandleStream(handle);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    if (handle->owner == WorkspaceOwner::Library && handle->workspace != nullptr) {
        handle->owned_size_hint = handle->workspace_size;
        ReleaseOwnedArena(handle);
    }
    handle->workspace = workspace;
    handle->workspace_size = workspaceSize;
    handle->owner = WorkspaceOwner::User;
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasSetLaunchEvents(aclblasHandle_t handle, void* const events[], int count)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (count < 0 || (count > 0 && events == nullptr)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    handle->launch_events.assign(static_cast<size_t>(count), nullptr);
    for (int i = 0; i < count; ++i) {
        handle->launch_events[static_cast<size_t>(i)] = static_cast<hipEvent_t>(events[i]);
    }
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasGetVersion(aclblasHandle_t handle, int* version)
{
    (void)handle;
    if (version == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    *version = kVersion;
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasLoggerConfigure(int logIsOn, int logToStdOut, int logToStdErr, const char* logFile)
{
    LoggerState& s = Logger();
    std::lock_guard<std::mutex> lock(s.mu);
    s.env_read = true;
    s.on = logIsOn != 0;
    s.to_stdout = logToStdOut != 0;
    s.to_stderr = logToStdErr != 0;
    s.file = logFile;
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasSetLoggerCallback(aclblasLogCallback userCallback)
{
    LoggerState& s = Logger();
    std::lock_guard<std::mutex> lock(s.mu);
    s.callback = userCallback;
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t aclblasGetLoggerCallback(aclblasLogCallback* userCallback)
{
    if (userCallback == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    LoggerState& s = Logger();
    std::lock_guard<std::mutex> lock(s.mu);
    *userCallback = s.callback;
    return ACLBLAS_STATUS_SUCCESS;
}

}  // extern "C"
