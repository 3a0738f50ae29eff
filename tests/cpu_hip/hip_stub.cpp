// A CPU model of the HIP runtime calls that libchemsim_hip.so's host code
// makes, for running that host code under AddressSanitizer and UBSan on a
// machine without a GPU.  Test infrastructure only.
//
// "Device" memory is host heap memory of exactly the requested size
// (posix_memalign, 256-byte aligned), so every copy, set and free the library
// issues is bounds-checked by ASan on both sides.  Kernel launches do nothing
// (device memory starts zeroed; values are meaningless, only memory safety is
// checked).  Streams are deferred queues: an asynchronous copy or memset runs
// only when something forces it -- a synchronisation of its stream or of an
// event recorded after it, a stream waiting on such an event, a blocking copy,
// a free, or the stream's destruction.  So a host buffer that is released
// before the copy that reads or writes it has been waited for is reported as
// a heap-use-after-free, which the real runtime may instead turn into silent
// heap corruption.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <vector>

namespace {

struct StubStream {
    std::deque<std::function<void()>> q;
    std::vector<StubStream*> waits;  // streams this one waits on (events)
};

struct StubEvent {
    StubStream* stream = nullptr;
};

std::recursive_mutex g_mu;
StubStream g_null;                      // the null stream
std::set<StubStream*> g_streams{&g_null};
std::set<void*> g_device, g_host;       // live allocations
bool g_capturing = false;
// Registered __device__ variables (host shadow -> bytes) and their "device"
// memory.  Constructed on first use: registration runs from the library's
// static constructors, possibly before this file's globals exist.
std::map<const void*, size_t>& var_size()
{
    static std::map<const void*, size_t> m;
    return m;
}
std::map<const void*, void*>& var_mem()
{
    static std::map<const void*, void*> m;
    return m;
}

struct CallConfig {
    dim3 grid, block;
    size_t shared;
    hipStream_t stream;
};
std::vector<CallConfig> g_configs;

StubStream* S(hipStream_t s) { return s ? reinterpret_cast<StubStream*>(s) : &g_null; }

void flush(StubStream* s)
{
    std::vector<StubStream*> deps;
    deps.swap(s->waits);
    for (StubStream* d : deps)
        if (g_streams.count(d)) flush(d);
    while (!s->q.empty()) {
        auto f = std::move(s->q.front());
        s->q.pop_front();
        f();
    }
}

void flush_all()
{
    const std::vector<StubStream*> all(g_streams.begin(), g_streams.end());
    for (StubStream* s : all) flush(s);
}

void enqueue(hipStream_t s, std::function<void()> f)
{
    if (g_capturing) return;  // captured work runs when the graph is launched (a no-op here)
    S(s)->q.push_back(std::move(f));
}

[[noreturn]] void die(const char* what, const void* p)
{
    std::fprintf(stderr, "hip_stub: %s %p\n", what, p);
    std::abort();
}

}  // namespace

extern "C" {

hipError_t hipGetDeviceCount(int* count)
{
    *count = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* prop, int)
{
    std::memset(prop, 0, sizeof *prop);
    std::snprintf(prop->name, sizeof prop->name, "cpu stub");
    std::snprintf(prop->gcnArchName, sizeof prop->gcnArchName, "gfx950");
    prop->multiProcessorCount = 256;
    prop->warpSize = 64;
    prop->maxThreadsPerBlock = 1024;
    prop->sharedMemPerBlock = 160 * 1024;
    return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "hip stub error"; }

hipError_t hipMalloc(void** ptr, size_t size)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    void* p = nullptr;
    if (posix_memalign(&p, 256, size ? size : 1) != 0) return hipErrorOutOfMemory;
    std::memset(p, 0, size);
    g_device.insert(p);
    *ptr = p;
    return hipSuccess;
}
hipError_t hipFree(void* p)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!p) return hipSuccess;
    flush_all();  // hipFree waits for the device
    if (!g_device.erase(p)) die("hipFree of a pointer hipMalloc did not return:", p);
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    void* p = nullptr;
    if (posix_memalign(&p, 256, size ? size : 1) != 0) return hipErrorOutOfMemory;
    g_host.insert(p);
    *ptr = p;
    return hipSuccess;
}
hipError_t hipHostFree(void* p)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!p) return hipSuccess;
    flush_all();
    if (!g_host.erase(p)) die("hipHostFree of a pointer hipHostMalloc did not return:", p);
    std::free(p);
    return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    flush_all();  // a blocking copy on the null stream waits for prior work
    if (n) std::memcpy(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t s)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    enqueue(s, [=] {
        if (n) std::memcpy(dst, src, n);
    });
    return hipSuccess;
}
hipError_t hipMemcpy2DAsync(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                            size_t height, hipMemcpyKind, hipStream_t s)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (width > dpitch || width > spitch) return hipErrorInvalidValue;
    enqueue(s, [=] {
        for (size_t r = 0; r < height; ++r)
            std::memcpy((char*)dst + r * dpitch, (const char*)src + r * spitch, width);
    });
    return hipSuccess;
}
hipError_t hipMemset(void* dst, int v, size_t n)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    flush_all();
    std::memset(dst, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int v, size_t n, hipStream_t s)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    enqueue(s, [=] { std::memset(dst, v, n); });
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    auto* st = new StubStream;
    g_streams.insert(st);
    *s = reinterpret_cast<hipStream_t>(st);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    StubStream* st = S(s);
    if (st == &g_null || !g_streams.count(st)) die("hipStreamDestroy of an unknown stream", s);
    flush(st);
    g_streams.erase(st);
    for (StubStream* o : g_streams) {
        auto& w = o->waits;
        for (size_t i = 0; i < w.size();)
            if (w[i] == st) w.erase(w.begin() + (long)i);
            else ++i;
    }
    delete st;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!g_streams.count(S(s))) die("hipStreamSynchronize of an unknown stream", s);
    flush(S(s));
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    auto* ev = reinterpret_cast<StubEvent*>(e);
    if (ev->stream && ev->stream != S(s)) S(s)->waits.push_back(ev->stream);
    return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = reinterpret_cast<hipEvent_t>(new StubEvent);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete reinterpret_cast<StubEvent*>(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    reinterpret_cast<StubEvent*>(e)->stream = S(s);
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    auto* ev = reinterpret_cast<StubEvent*>(e);
    if (ev->stream && g_streams.count(ev->stream)) flush(ev->stream);
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.001f;
    return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    g_capturing = true;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g)
{
    g_capturing = false;
    *g = reinterpret_cast<hipGraph_t>(new char[1]);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* ge, hipGraph_t, hipGraphNode_t*, char*, size_t)
{
    *ge = reinterpret_cast<hipGraphExec_t>(new char[1]);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete[] reinterpret_cast<char*>(g);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t g)
{
    delete[] reinterpret_cast<char*>(g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }

// Kernel launches: the configuration stack of <<<>>> and hipLaunchKernel.
hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shared, hipStream_t s)
{
    g_configs.push_back({grid, block, shared, s});
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* grid, dim3* block, size_t* shared, hipStream_t* s)
{
    if (g_configs.empty()) return hipErrorInvalidConfiguration;
    const CallConfig c = g_configs.back();
    g_configs.pop_back();
    *grid = c.grid;
    *block = c.block;
    *shared = c.shared;
    *s = c.stream;
    return hipSuccess;
}
hipError_t hipLaunchKernel(const void*, dim3 grid, dim3 block, void**, size_t shared, hipStream_t)
{
    if (grid.x == 0 || grid.y == 0 || grid.z == 0 || block.x * block.y * block.z == 0 ||
        block.x * block.y * block.z > 1024 || shared > 160 * 1024)
        return hipErrorInvalidConfiguration;
    return hipSuccess;
}
// Cooperative launches (a resident grid) check the same configuration; one
// 1024-thread block per CU of the model device's 256 fits.
hipError_t hipLaunchCooperativeKernel(const void* f, dim3 grid, dim3 block, void** args, unsigned int shared,
                                      hipStream_t s)
{
    if (grid.x * grid.y * grid.z > 256u * (1024u / std::max(1u, block.x * block.y * block.z)))
        return hipErrorCooperativeLaunchTooLarge;
    return hipLaunchKernel(f, grid, block, args, shared, s);
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int block, size_t)
{
    if (!n || block <= 0 || block > 1024) return hipErrorInvalidValue;
    *n = 1024 / block;
    return hipSuccess;
}
void** __hipRegisterFatBinary(const void*)
{
    static void* module = nullptr;
    return &module;
}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, void*, void*, void*, void*, int*) {}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterVar(void**, void* var, char*, char*, int, size_t size, int, int)
{
    var_size()[var] = size;
}
// A __device__ variable's address: exactly-sized zeroed heap memory per
// variable, kept until the process ends (as a code object's globals are).
hipError_t hipGetSymbolAddress(void** ptr, const void* symbol)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    auto it = var_size().find(symbol);
    if (it == var_size().end()) return hipErrorInvalidSymbol;
    void*& mem = var_mem()[symbol];
    if (!mem) {
        if (posix_memalign(&mem, 256, it->second ? it->second : 1) != 0) return hipErrorOutOfMemory;
        std::memset(mem, 0, it->second);
    }
    *ptr = mem;
    return hipSuccess;
}

}  // extern "C"
