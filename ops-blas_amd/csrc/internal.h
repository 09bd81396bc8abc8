// Library-internal declarations shared by the host translation units.
//
// The handle keeps the reference's workspace policy (reference
// blas/common/helper/aclblas_handle_internal.h:26-209): one active workspace
// that is either library-owned (32 MiB default, grows by doubling up to 2 GiB)
// or borrowed from the caller (grow-only, never freed by the library).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "cann_ops_blas.h"

namespace aclblas {

constexpr size_t kDefaultWorkspaceSize = 32u * 1024u * 1024u;
constexpr size_t kMaxWorkspaceSize = 2ull * 1024ull * 1024ull * 1024ull;
// MAJOR*10000 + MINOR*100 + PATCH (reference blas/common/helper/aclblas_version.h:18-31)
constexpr int kVersion = 10000;
// Largest matrix order of the batched LAPACK calls (row indices are packed in
// 16 bits by the pivot search).
constexpr int kMaxLapackN = 65535;

enum class WorkspaceOwner { Library, User };

}  // namespace aclblas

struct _aclblas_handle {
    hipStream_t stream = nullptr;
    void* workspace = nullptr;
    size_t workspace_size = 0;
    aclblas::WorkspaceOwner owner = aclblas::WorkspaceOwner::Library;
    // Size of the handle's own arena to restore after a borrowed workspace.
    size_t owned_size_hint = 0;
    // Set when a call captured into a graph used the handle's own arena; such
    // an arena is parked (freed by aclblasDestroy) rather than freed when it is
    // replaced.  Graphs captured on a handle must not be replayed after
    // aclblasDestroy, nor after a borrowed workspace they used is freed.
    bool arena_in_graph = false;
    std::vector<void*> parked_arenas;
    // Device the handle was created on and its compute-unit count (used to size
    // persistent grids).  device < 0 means no GPU was visible at creation.
    int device = -1;
    int cu_count = 0;
    // Pinned host staging for per-call parameter tables (grouped GEMM), copied
    // to the workspace with an asynchronous copy; stage_event marks the end of
    // the last such copy so the staging memory is not rewritten under it.
    void* host_stage = nullptr;
    size_t host_stage_size = 0;
    hipEvent_t stage_event = nullptr;
    // Pinned copies of tables staged while the stream was being captured into
    // a graph (each replay reads its own copy); freed with the handle.
    std::vector<void*> captured_stages;
    // Optional events recorded around the kernels of multi-launch calls
    // (aclblasSetLaunchEvents): event 2i before and 2i+1 after launch i.
    std::vector<hipEvent_t> launch_events;
    // Per-operation forced launch choice {set, a, b} (aclblasSetTuningOverride);
    // set == 0 means the measured table decides.
    int tune_override[4][3] = {};
};

namespace aclblas {

// Logging: host-side only, off by default, configured through
// aclblasLoggerConfigure / aclblasSetLoggerCallback or ACLBLAS_LOG_LEVEL.
void LogError(const char* api, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void LogInfo(const char* api, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Synchronises the handle's stream; EXECUTION_FAILED on error.
aclblasStatus_t SyncHandleStream(_aclblas_handle* h);

// Makes the active workspace at least `bytes` long, growing the handle's own
// arena by doubling when needed (reference aclblas_handle_internal.h:162-209);
// a borrowed workspace that is too small gives ALLOC_FAILED.  Every call that
// addresses the workspace goes through here or through MarkWorkspaceUse.
aclblasStatus_t EnsureWorkspace(_aclblas_handle* h, size_t bytes);

// Records that the current call addresses the active workspace as it is (no
// growth); inside a stream capture this keeps the arena alive for the graph.
void MarkWorkspaceUse(_aclblas_handle* h);

// True while the handle's stream is being captured into a graph.
bool StreamCapturing(const _aclblas_handle* h);

// Copies `bytes` of host data to the start of the handle's workspace,
// asynchronously on the handle's stream, through the handle's pinned staging
// buffer; *dev receives the device address.  Blocks only while a previous
// staging copy of this handle is still pending.
aclblasStatus_t StageToWorkspace(_aclblas_handle* h, const void* src, size_t bytes, void** dev);

// Records the handle's launch event `slot` (aclblasSetLaunchEvents) on its
// stream, if the caller registered one.
inline void MarkLaunch(_aclblas_handle* h, size_t slot)
{
    if (slot < h->launch_events.size() && h->launch_events[slot] != nullptr) {
        (void)hipEventRecord(h->launch_events[slot], h->stream);
    }
}

// Checks that the handle can launch work and that the last launch was accepted.
inline bool HandleHasDevice(const _aclblas_handle* h) { return h->device >= 0; }

inline aclblasStatus_t LaunchStatus(const char* api)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        LogError(api, "kernel launch failed: %s", hipGetErrorString(e));
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

// C callers pass enumeration arguments as plain ints.  Reading a value outside
// an enumeration's range through the C++ enum type is undefined behaviour, so
// the validators take enum arguments by reference and read their bits as an
// int; code after validation uses the enum types normally.
template <typename E>
inline int EnumBits(const E& e)
{
    static_assert(sizeof(E) == sizeof(int), "C enumerations are int-sized");
    int v;
    std::memcpy(&v, &e, sizeof v);
    return v;
}

inline bool IsValidOp(const aclblasOperation_t& op)
{
    const int v = EnumBits(op);
    return v == ACLBLAS_OP_N || v == ACLBLAS_OP_T || v == ACLBLAS_OP_C;
}

inline int MaxInt(int a, int b) { return a > b ? a : b; }

}  // namespace aclblas
