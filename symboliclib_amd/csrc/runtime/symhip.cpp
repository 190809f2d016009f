// MI355X runtime for compiled expression programs: device buffers, hipRTC compilation with
// an on-disk code-object cache, and ordered kernel launches.  C ABI declared in symhip.h.
#include "symhip.h"

#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../sym/Hashing.hpp"
#include "symhip_internal.hpp"

namespace {

thread_local std::string g_lastError;

int fail(int code, const std::string& msg) {
    g_lastError = msg;
    return code;
}

int hipFail(hipError_t e, const char* what) {
    g_lastError = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e;
}

#define HIP_TRY(call, what)                         \
    do {                                            \
        hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return hipFail(e_, what); \
    } while (0)

// A captured graph holds the launch arguments: drop it when they change.
void dropGraph(symhip_t h) {
    if (h->graphExec) {
        (void)hipGraphExecDestroy(h->graphExec);
        h->graphExec = nullptr;
    }
    if (h->graph) {
        (void)hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
}

std::string defaultArch() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return "gfx950";
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "gfx950";
    }
    std::string a = prop.gcnArchName;
    auto colon = a.find(':');
    if (colon != std::string::npos) a = a.substr(0, colon);
    return a.empty() ? std::string("gfx950") : a;
}

std::vector<std::string> splitOptions(const char* options) {
    std::vector<std::string> out;
    if (!options) return out;
    std::istringstream ss(options);
    std::string tok;
    while (ss >> tok) out.push_back(tok);
    return out;
}

// Directory of the code-object cache: $SYMLIB_CACHE_DIR, else .codecache next to the
// directory holding this library.
std::string cacheDir() {
    if (const char* env = std::getenv("SYMLIB_CACHE_DIR")) return env;
    Dl_info info;
    if (dladdr((void*)&cacheDir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        auto slash = p.rfind('/');
        if (slash != std::string::npos) {
            std::string dir = p.substr(0, slash);  // .../symboliclib_amd/lib
            auto up = dir.rfind('/');
            if (up != std::string::npos) return dir.substr(0, up) + "/.codecache";
        }
    }
    return ".codecache";
}

// Every option hipRTC compiles a program with: the fixed ones, then the caller's extras.
std::vector<std::string> compileOptions(const std::vector<std::string>& extra, const std::string& arch) {
    std::vector<std::string> opts = {"--offload-arch=" + arch, "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                     "-Xclang", "-packed-fp32-ops"};
    opts.insert(opts.end(), extra.begin(), extra.end());
    return opts;
}

int compileSource(const std::string& src, const std::vector<std::string>& extra, const std::string& arch,
                  std::vector<char>& code) {
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "symlib_program.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) return fail(SYMHIP_E_COMPILE, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));

    const std::vector<std::string> opts = compileOptions(extra, arch);
    std::vector<const char*> copts;
    for (auto& o : opts) copts.push_back(o.c_str());

    r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    size_t logSize = 0;
    hiprtcGetProgramLogSize(prog, &logSize);
    std::string log(logSize, '\0');
    if (logSize) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(SYMHIP_E_COMPILE, std::string("hipRTC compilation failed: ") + hiprtcGetErrorString(r) + "\n" + log);
    }
    size_t codeSize = 0;
    hiprtcGetCodeSize(prog, &codeSize);
    code.resize(codeSize);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return SYMHIP_SUCCESS;
}

// Cache key of a code object: hash of the source, the full option list it is compiled with
// and the hipRTC version, so that a changed flag or compiler never reuses an old object.
std::string cacheKey(const std::string& src, const char* options, const std::string& arch) {
    int major = 0, minor = 0;
    if (hiprtcVersion(&major, &minor) != HIPRTC_SUCCESS) major = minor = -1;
    std::string all = src + "\n//opts:";
    for (const auto& o : compileOptions(splitOptions(options), arch)) all += o + " ";
    all += "\n//hiprtc:" + std::to_string(major) + "." + std::to_string(minor);
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%016llx", (unsigned long long)Sym::CRC64_hash(all.data(), all.size()));
    return std::string(buf) + "_" + std::to_string(all.size());
}

bool readFile(const std::string& path, std::vector<char>& data) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    data.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !data.empty();
}

bool writeFileAtomic(const std::string& path, const std::vector<char>& data) {
    static std::atomic<unsigned> serial{0};
    std::string tmp = path + ".tmp" + std::to_string(getpid()) + "_" + std::to_string(serial++);
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return false;
        f.write(data.data(), (std::streamsize)data.size());
        if (!f) return false;
    }
    return std::rename(tmp.c_str(), path.c_str()) == 0;
}

}  // namespace

extern "C" {

int symhip_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n > 0 ? 1 : 0;
}

const char* symhip_last_error(void) { return g_lastError.c_str(); }

void symhip_destroy(symhip_t h) {
    if (!h) return;
    if (h->graphExec) (void)hipGraphExecDestroy(h->graphExec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->module) (void)hipModuleUnload(h->module);
    // a destroy cannot report errors: the buffers are released whatever the frees return
    for (void* ptr : {(void*)h->x, (void*)h->result, (void*)h->c, (void*)h->p, (void*)h->o, (void*)h->tiles, (void*)h->work})
        if (ptr) (void)hipFree(ptr);
    delete h;
}

int symhip_create(symhip_t* out, const symhip_program_desc* d) {
    if (!out || !d) return fail(SYMHIP_E_INVALID, "symhip_create: null argument");
    *out = nullptr;
    if (!symhip_device_available()) return fail(SYMHIP_E_NO_DEVICE, "symhip_create: no HIP device visible");

    auto* h = new symhip_program();
    h->numValues = d->num_values;
    h->argStart.assign(d->arg_start, d->arg_start + d->num_args);
    h->argLength.assign(d->arg_length, d->arg_length + d->num_args);
    h->outStart.assign(d->out_start, d->out_start + d->num_outputs);
    h->outLength.assign(d->out_length, d->out_length + d->num_outputs);
    for (uint32_t i = 0; i < d->num_args; ++i)
        if (h->argStart[i] + h->argLength[i] > h->numValues) {
            delete h;
            return fail(SYMHIP_E_INVALID, "symhip_create: input container outside the value buffer");
        }
    for (uint32_t k = 0; k < d->num_outputs; ++k) h->numResults = std::max(h->numResults, h->outStart[k] + h->outLength[k]);
    for (uint32_t i = 0; i < d->num_launches; ++i) {
        h->names.push_back(d->launches[i].name ? d->launches[i].name : "");
        h->launches.push_back(d->launches[i]);
    }
    h->refForm.assign(h->launches.size(), 0);
    h->refOffsets.assign(h->launches.size(), {0, 0, 0});
    h->numConst = d->num_const;
    h->numIndex = d->num_index;
    h->numOutpos = d->num_outpos;

    auto upload = [&](void** dst, const void* src, size_t bytes, const char* what) -> int {
        if (bytes == 0) return SYMHIP_SUCCESS;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return hipFail(e, what);
        if (src) {
            e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) return hipFail(e, what);
        } else {
            e = hipMemset(*dst, 0, bytes);
            if (e != hipSuccess) return hipFail(e, what);
        }
        return SYMHIP_SUCCESS;
    };
    int rc;
    if ((rc = upload((void**)&h->x, nullptr, sizeof(double) * h->numValues, "hipMalloc(values)")) ||
        (rc = upload((void**)&h->result, nullptr, sizeof(double) * h->numResults, "hipMalloc(results)")) ||
        (rc = upload((void**)&h->p, d->index_data, sizeof(uint32_t) * d->num_index, "upload(index)")) ||
        (rc = upload((void**)&h->c, d->const_data, sizeof(double) * d->num_const, "upload(consts)")) ||
        (rc = upload((void**)&h->o, d->outpos_data, sizeof(uint32_t) * d->num_outpos, "upload(outpos)")) ||
        (rc = upload((void**)&h->tiles, d->tiles_data, sizeof(uint32_t) * d->num_tiles_data, "upload(tiles)")) ||
        (rc = upload((void**)&h->work, nullptr, sizeof(uint32_t) * d->num_work_words, "hipMalloc(work)"))) {
        symhip_destroy(h);
        return rc;
    }
    h->numTilesData = d->num_tiles_data;
    h->numWork = d->num_work_words;
    *out = h;
    return SYMHIP_SUCCESS;
}

int symhip_compile_to_file(const char* source, const char* options, const char* arch, const char* out_path) {
    if (!source || !out_path) return fail(SYMHIP_E_INVALID, "symhip_compile_to_file: null argument");
    std::vector<char> code;
    int rc = compileSource(source, splitOptions(options), arch ? arch : "gfx950", code);
    if (rc) return rc;
    if (!writeFileAtomic(out_path, code)) return fail(SYMHIP_E_IO, std::string("cannot write ") + out_path);
    return SYMHIP_SUCCESS;
}

int symhip_cache_compile(const char* source, const char* options, const char* arch) {
    if (!source) return fail(SYMHIP_E_INVALID, "symhip_cache_compile: null source");
    const std::string a = arch ? std::string(arch) : defaultArch();
    const std::string dir = cacheDir();
    const std::string path = dir + "/" + cacheKey(source, options, a) + "." + a + ".co";
    struct stat st;
    if (stat(path.c_str(), &st) == 0 && st.st_size > 0) return SYMHIP_SUCCESS;
    std::vector<char> code;
    int rc = compileSource(source, splitOptions(options), a, code);
    if (rc) return rc;
    mkdir(dir.c_str(), 0755);
    if (!writeFileAtomic(path, code)) return fail(SYMHIP_E_IO, "cannot write " + path);
    return SYMHIP_SUCCESS;
}

int symhip_compile(symhip_t h, const char* source, const char* options) {
    if (!h || !source) return fail(SYMHIP_E_INVALID, "symhip_compile: null argument");
    const std::string arch = defaultArch();
    const std::string src = source;
    const std::string dir = cacheDir();
    const std::string path = dir + "/" + cacheKey(src, options, arch) + "." + arch + ".co";

    std::vector<char> code;
    if (!readFile(path, code)) {
        int rc = compileSource(src, splitOptions(options), arch, code);
        if (rc) return rc;
        mkdir(dir.c_str(), 0755);
        writeFileAtomic(path, code);  // cache failures are not fatal
    }

    if (h->module) {
        (void)hipModuleUnload(h->module);
        h->module = nullptr;
    }
    h->functions.clear();
    HIP_TRY(hipModuleLoadData(&h->module, code.data()), "hipModuleLoadData");
    for (auto& name : h->names) {
        hipFunction_t f;
        hipError_t e = hipModuleGetFunction(&f, h->module, name.c_str());
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SYMHIP_E_SYMBOL, "symhip_compile: module has no function " + name);
        }
        h->functions.push_back(f);
    }
    h->grids.assign(h->launches.size(), 0);
    for (size_t i = 0; i < h->launches.size(); ++i) {
        const symhip_launch& L = h->launches[i];
        if (L.shared > 65536)
            HIP_TRY(hipFuncSetAttribute((const void*)h->functions[i], hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)L.shared),
                    "hipFuncSetAttribute(max dynamic LDS)");
        h->grids[i] = L.grid;
        if (L.grid == SYMHIP_GRID_FILL) {
            int perCU = 0, dev = 0, cus = 0;
            HIP_TRY(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, h->functions[i], (int)L.block, L.shared),
                    "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
            HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
            HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
            if (perCU <= 0) return fail(SYMHIP_E_INVALID, "symhip_compile: " + h->names[i] + " cannot be resident (block or LDS too large)");
            h->grids[i] = (uint32_t)perCU * (uint32_t)cus;
        }
    }
    if (h->graphExec) {
        (void)hipGraphExecDestroy(h->graphExec);
        h->graphExec = nullptr;
    }
    if (h->graph) {
        (void)hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
    return SYMHIP_SUCCESS;
}

int symhip_set_arg(symhip_t h, int i, const double* host, size_t n) {
    if (!h || i < 0 || (size_t)i >= h->argStart.size()) return fail(SYMHIP_E_INVALID, "symhip_set_arg: bad argument index");
    if (n != h->argLength[i]) return fail(SYMHIP_E_INVALID, "symhip_set_arg: length does not match the input container");
    if (!n) return SYMHIP_SUCCESS;
    if (!host) return fail(SYMHIP_E_INVALID, "symhip_set_arg: null data");
    // launches of this program may still run on any stream (a non-blocking one included) and
    // read the inputs: wait for them before the copy overwrites the inputs
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize(setArg)");
    HIP_TRY(hipMemcpy(h->x + h->argStart[i], host, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy(setArg)");
    return SYMHIP_SUCCESS;
}

int symhip_set_arg_async(symhip_t h, int i, const double* host, size_t n, void* stream) {
    if (!h || i < 0 || (size_t)i >= h->argStart.size()) return fail(SYMHIP_E_INVALID, "symhip_set_arg_async: bad argument index");
    if (n != h->argLength[i]) return fail(SYMHIP_E_INVALID, "symhip_set_arg_async: length does not match the input container");
    if (!n) return SYMHIP_SUCCESS;
    if (!host) return fail(SYMHIP_E_INVALID, "symhip_set_arg_async: null data");
    HIP_TRY(hipMemcpyAsync(h->x + h->argStart[i], host, sizeof(double) * n, hipMemcpyHostToDevice, (hipStream_t)stream),
            "hipMemcpyAsync(setArg)");
    return SYMHIP_SUCCESS;
}

int symhip_set_arg_device(symhip_t h, int i, const double* dev, size_t n, void* stream) {
    if (!h || i < 0 || (size_t)i >= h->argStart.size()) return fail(SYMHIP_E_INVALID, "symhip_set_arg_device: bad argument index");
    if (n != h->argLength[i]) return fail(SYMHIP_E_INVALID, "symhip_set_arg_device: length mismatch");
    if (!n) return SYMHIP_SUCCESS;
    HIP_TRY(hipMemcpyAsync(h->x + h->argStart[i], dev, sizeof(double) * n, hipMemcpyDeviceToDevice, (hipStream_t)stream),
            "hipMemcpyAsync(setArg)");
    return SYMHIP_SUCCESS;
}

uint32_t symhip_num_launches(symhip_t h) { return h ? (uint32_t)h->launches.size() : 0; }

uint32_t symhip_launch_grid(symhip_t h, uint32_t idx) {
    if (!h || idx >= h->launches.size()) return 0;
    return idx < h->grids.size() ? h->grids[idx] : h->launches[idx].grid;
}

int symhip_launch_one(symhip_t h, uint32_t idx, void* stream) {
    if (!h) return fail(SYMHIP_E_INVALID, "symhip_launch_one: null handle");
    return symhipLaunch(h, idx, stream, h->work);
}

int symhip_run_graph(symhip_t h, void* stream) {
    if (!h) return fail(SYMHIP_E_INVALID, "symhip_run_graph: null handle");
    if (h->functions.size() != h->launches.size()) return fail(SYMHIP_E_NOT_COMPILED, "symhip: program not compiled");
    if (!h->graphExec) {
        hipStream_t cs = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        int rc = SYMHIP_SUCCESS;
        if (e == hipSuccess) {
            rc = symhipLaunchAll(h, cs, h->work);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(cs, &g);
            if (!rc && e2 != hipSuccess) rc = hipFail(e2, "hipStreamEndCapture");
            if (!rc) {
                e2 = hipGraphInstantiate(&h->graphExec, g, nullptr, nullptr, 0);
                if (e2 != hipSuccess) rc = hipFail(e2, "hipGraphInstantiate");
            }
            if (!rc) h->graph = g;
            else if (g) (void)hipGraphDestroy(g);
        } else {
            rc = hipFail(e, "hipStreamBeginCapture");
        }
        (void)hipStreamDestroy(cs);
        if (rc) {
            h->graphExec = nullptr;
            return rc;
        }
    }
    HIP_TRY(hipGraphLaunch(h->graphExec, (hipStream_t)stream), "hipGraphLaunch");
    return SYMHIP_SUCCESS;
}

int symhip_run(symhip_t h, void* stream) {
    if (!h) return fail(SYMHIP_E_INVALID, "symhip_run: null handle");
    return symhipLaunchAll(h, stream, h->work);
}

int symhip_set_launch_offsets(symhip_t h, uint32_t idx, uint64_t c_offset, uint64_t y_offset, uint64_t p_offset) {
    if (!h || idx >= h->launches.size()) return fail(SYMHIP_E_INVALID, "symhip_set_launch_offsets: bad launch index");
    if (c_offset > h->numConst || y_offset > h->numValues || p_offset > h->numIndex)
        return fail(SYMHIP_E_INVALID, "symhip_set_launch_offsets: offset beyond its table (c " + std::to_string(c_offset) +
                                          " of " + std::to_string(h->numConst) + ", y " + std::to_string(y_offset) +
                                          " of " + std::to_string(h->numValues) + ", p " + std::to_string(p_offset) +
                                          " of " + std::to_string(h->numIndex) + ")");
    h->refForm[idx] = 1;
    h->refOffsets[idx] = {c_offset, y_offset, p_offset};
    dropGraph(h);
    return SYMHIP_SUCCESS;
}

const char* symhip_reference_prelude(void) {
    return "#define KERNEL extern \"C\" __global__\n"
           "typedef unsigned int INDEX;\n"
           "typedef double RealT;\n"
           "#define restrict __restrict__\n"
           "#define hipThreadIdx_x threadIdx.x\n"
           "#define hipBlockDim_x blockDim.x\n"
           "#define hipBlockIdx_x blockIdx.x\n"
           "#define hipGridDim_x gridDim.x\n";
}

int symhip_set_result_gather(symhip_t h, int on) {
    if (!h) return fail(SYMHIP_E_INVALID, "symhip_set_result_gather: null handle");
    if (on) {
        if (h->numOutpos != h->numResults)
            return fail(SYMHIP_E_INVALID, "symhip_set_result_gather: the placement table has " +
                                              std::to_string(h->numOutpos) + " entries for " +
                                              std::to_string(h->numResults) + " results");
        // every gather position must lie inside x
        std::vector<uint32_t> pos(h->numOutpos);
        if (!pos.empty())
            HIP_TRY(hipMemcpy(pos.data(), h->o, sizeof(uint32_t) * pos.size(), hipMemcpyDeviceToHost), "hipMemcpy(outpos)");
        for (uint32_t q : pos)
            if (q >= h->numValues)
                return fail(SYMHIP_E_INVALID, "symhip_set_result_gather: position " + std::to_string(q) +
                                                  " outside the value buffer of " + std::to_string(h->numValues));
    }
    h->gather = on != 0;
    dropGraph(h);
    return SYMHIP_SUCCESS;
}

int symhip_update_tiles(symhip_t h, uint64_t offset, uint64_t n, const uint32_t* data) {
    if (!h || (n && !data) || offset + n > h->numTilesData) return fail(SYMHIP_E_INVALID, "symhip_update_tiles: range outside the tile pool");
    if (!n) return SYMHIP_SUCCESS;
    // a launch still running on any stream may be reading the pool
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize(update_tiles)");
    HIP_TRY(hipMemcpy(h->tiles + offset, data, sizeof(uint32_t) * n, hipMemcpyHostToDevice), "hipMemcpy(tiles)");
    return SYMHIP_SUCCESS;
}

int symhip_set_launch(symhip_t h, uint32_t idx, const symhip_launch* launch) {
    if (!h || !launch || idx >= h->launches.size()) return fail(SYMHIP_E_INVALID, "symhip_set_launch: bad launch index");
    h->names[idx] = launch->name ? launch->name : "";
    h->launches[idx] = *launch;
    h->functions.clear();  // compile again before running
    if (h->graphExec) {
        (void)hipGraphExecDestroy(h->graphExec);
        h->graphExec = nullptr;
    }
    if (h->graph) {
        (void)hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
    return SYMHIP_SUCCESS;
}

int symhip_set_launches(symhip_t h, uint32_t n, const symhip_launch* launches) {
    if (!h || (n && !launches)) return fail(SYMHIP_E_INVALID, "symhip_set_launches: null argument");
    h->names.clear();
    h->launches.clear();
    for (uint32_t i = 0; i < n; ++i) {
        h->names.push_back(launches[i].name ? launches[i].name : "");
        h->launches.push_back(launches[i]);
    }
    h->refForm.assign(n, 0);
    h->refOffsets.assign(n, {0, 0, 0});
    h->functions.clear();  // compile again before running
    h->grids.clear();
    if (h->graphExec) {
        (void)hipGraphExecDestroy(h->graphExec);
        h->graphExec = nullptr;
    }
    if (h->graph) {
        (void)hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
    return SYMHIP_SUCCESS;
}

int symhip_set_work_words(symhip_t h, uint64_t n) {
    if (!h) return fail(SYMHIP_E_INVALID, "symhip_set_work_words: null handle");
    if (n == h->numWork) return SYMHIP_SUCCESS;
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_TRY(hipFree(h->work), "hipFree(work)");
    h->work = nullptr;
    h->numWork = 0;
    if (n) {
        HIP_TRY(hipMalloc((void**)&h->work, sizeof(uint32_t) * n), "hipMalloc(work)");
        HIP_TRY(hipMemset(h->work, 0, sizeof(uint32_t) * n), "hipMemset(work)");
    }
    h->numWork = n;
    if (h->graphExec) {
        (void)hipGraphExecDestroy(h->graphExec);
        h->graphExec = nullptr;
    }
    if (h->graph) {
        (void)hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
    return SYMHIP_SUCCESS;
}

int symhip_read_work(symhip_t h, uint64_t offset, uint64_t n, uint32_t* out) {
    if (!h || (n && !out) || offset + n > h->numWork) return fail(SYMHIP_E_INVALID, "symhip_read_work: range outside the work words");
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (n) HIP_TRY(hipMemcpy(out, h->work + offset, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "hipMemcpy(work)");
    return SYMHIP_SUCCESS;
}

int symhip_launch_info(symhip_t h, uint32_t idx, int* vgprs, int* blocks_per_cu, int* cus) {
    if (!h || idx >= h->launches.size()) return fail(SYMHIP_E_INVALID, "symhip_launch_info: bad launch index");
    if (h->functions.size() != h->launches.size()) return fail(SYMHIP_E_NOT_COMPILED, "symhip: program not compiled");
    int dev = 0, n = 0, per = 0, regs = 0;
    HIP_TRY(hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, h->functions[idx]), "hipFuncGetAttribute");
    HIP_TRY(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per, h->functions[idx], (int)h->launches[idx].block,
                                                               h->launches[idx].shared),
            "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    HIP_TRY(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    if (vgprs) *vgprs = regs;
    if (blocks_per_cu) *blocks_per_cu = per;
    if (cus) *cus = n;
    return SYMHIP_SUCCESS;
}

int symhip_wall_clock_khz(int* khz) {
    if (!khz) return fail(SYMHIP_E_INVALID, "symhip_wall_clock_khz: null argument");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    HIP_TRY(hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
    return SYMHIP_SUCCESS;
}

int symhip_get_result(symhip_t h, int k, double* host, size_t n) {
    if (!h || k < 0 || (size_t)k >= h->outStart.size()) return fail(SYMHIP_E_INVALID, "symhip_get_result: bad output index");
    if (n != h->outLength[k]) return fail(SYMHIP_E_INVALID, "symhip_get_result: length does not match the output container");
    if (!n) return SYMHIP_SUCCESS;
    if (!host) return fail(SYMHIP_E_INVALID, "symhip_get_result: null data");
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_TRY(hipMemcpy(host, h->result + h->outStart[k], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy(getResult)");
    return SYMHIP_SUCCESS;
}

int symhip_get_result_device(symhip_t h, int k, double* dev, size_t n, void* stream) {
    if (!h || k < 0 || (size_t)k >= h->outStart.size()) return fail(SYMHIP_E_INVALID, "symhip_get_result_device: bad output index");
    if (n != h->outLength[k]) return fail(SYMHIP_E_INVALID, "symhip_get_result_device: length mismatch");
    if (!n) return SYMHIP_SUCCESS;
    HIP_TRY(hipMemcpyAsync(dev, h->result + h->outStart[k], sizeof(double) * n, hipMemcpyDeviceToDevice, (hipStream_t)stream),
            "hipMemcpyAsync(getResult)");
    return SYMHIP_SUCCESS;
}

const double* symhip_output_device_ptr(symhip_t h, int k) {
    if (!h || k < 0 || (size_t)k >= h->outStart.size()) return nullptr;
    return h->result + h->outStart[k];
}

double* symhip_values_device_ptr(symhip_t h) { return h ? h->x : nullptr; }

}  // extern "C"

int symhipLaunch(symhip_t h, uint32_t idx, void* stream, uint32_t* work) {
    if (!h || idx >= h->launches.size()) return fail(SYMHIP_E_INVALID, "symhip_launch_one: bad launch index");
    if (h->functions.size() != h->launches.size()) return fail(SYMHIP_E_NOT_COMPILED, "symhip: program not compiled");
    const symhip_launch& L = h->launches[idx];
    const uint32_t grid = h->grids[idx];
    if (grid == 0) return SYMHIP_SUCCESS;
    if (h->refForm[idx]) {
        // the reference's K(num): k<i>(x, c + constantsOffset[i], x + outputOffset[i], p + indexOffset[i])
        const auto& off = h->refOffsets[idx];
        double* c = h->c ? h->c + off[0] : nullptr;
        double* y = h->x ? h->x + off[1] : nullptr;
        uint32_t* p = h->p ? h->p + off[2] : nullptr;
        void* args[] = {&h->x, &c, &y, &p};
        HIP_TRY(hipModuleLaunchKernel(h->functions[idx], grid, 1, 1, L.block, 1, 1, L.shared, (hipStream_t)stream, args,
                                      nullptr),
                "hipModuleLaunchKernel");
        return SYMHIP_SUCCESS;
    }
    void* args[] = {&h->x, &h->x, &h->c, &h->p, &h->o, &h->result, &h->tiles, &work};
    HIP_TRY(hipModuleLaunchKernel(h->functions[idx], grid, 1, 1, L.block, 1, 1, L.shared, (hipStream_t)stream, args,
                                  nullptr),
            "hipModuleLaunchKernel");
    return SYMHIP_SUCCESS;
}

int symhipLaunchAll(symhip_t h, void* stream, uint32_t* work) {
    if (h->functions.size() != h->launches.size()) return fail(SYMHIP_E_NOT_COMPILED, "symhip: program not compiled");
    for (uint32_t i = 0; i < h->launches.size(); ++i) {
        const int rc = symhipLaunch(h, i, stream, work);
        if (rc) return rc;
    }
    return h->gather ? symhipGather(h, stream) : SYMHIP_SUCCESS;
}

int symhipFail(int code, const std::string& msg) { return fail(code, msg); }
