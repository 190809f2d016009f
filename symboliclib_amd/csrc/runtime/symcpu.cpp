// Host runtime of the CPU form (symcpu.h): compile the emitted host source with clang and the
// reference's flags into a shared library kept in the code cache, load it, and run it over
// host copies of the program's tables.
#include "symcpu.h"

#include <dlfcn.h>
#include <spawn.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../sym/Hashing.hpp"

extern char** environ;

namespace {

thread_local std::string g_cpuError;

int fail(int code, const std::string& msg) {
    g_cpuError = msg;
    return code;
}

typedef void (*RunFn)(const double*, double*, const double*, const uint32_t*, const uint32_t*, double*, int);

// Flags of the reference's CPU compile (src/ComputeUnit.cpp:125, :132), the thread count
// passed at run time instead of -DNUMTHREADS.
const char* kReferenceFlags =
    "-O3 -ffast-math -fno-trapping-math -fno-math-errno -fno-signed-zeros -std=c++17 -msse4.2 -mavx2 -mfma -fopenmp";

std::string compiler() {
    if (const char* e = std::getenv("SYMLIB_CPU_CXX")) return e;
    return "/opt/rocm/llvm/bin/clang++";
}

std::string cacheDirectory() {
    if (const char* env = std::getenv("SYMLIB_CACHE_DIR")) return env;
    Dl_info info;
    if (dladdr((void*)&cacheDirectory, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        auto slash = p.rfind('/');
        if (slash != std::string::npos) {
            std::string dir = p.substr(0, slash);
            auto up = dir.rfind('/');
            if (up != std::string::npos) return dir.substr(0, up) + "/.codecache";
        }
    }
    return ".codecache";
}

// Run argv as a child process (posix_spawn, no shell); returns its exit status (or -1) and
// its combined output.
int runProcess(const std::vector<std::string>& argv, std::string& output) {
    int pipefd[2];
    if (pipe(pipefd) != 0) return -1;
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_adddup2(&fa, pipefd[1], 1);
    posix_spawn_file_actions_adddup2(&fa, pipefd[1], 2);
    posix_spawn_file_actions_addclose(&fa, pipefd[0]);
    std::vector<char*> a;
    for (const auto& s : argv) a.push_back(const_cast<char*>(s.c_str()));
    a.push_back(nullptr);
    pid_t pid = 0;
    const int rc = posix_spawn(&pid, a[0], &fa, nullptr, a.data(), environ);
    posix_spawn_file_actions_destroy(&fa);
    close(pipefd[1]);
    if (rc != 0) {
        close(pipefd[0]);
        output = std::string("cannot start ") + a[0] + ": " + std::strerror(rc);
        return -1;
    }
    char buf[4096];
    ssize_t r;
    while ((r = read(pipefd[0], buf, sizeof(buf))) > 0) output.append(buf, (size_t)r);
    close(pipefd[0]);
    int status = 0;
    if (waitpid(pid, &status, 0) < 0) return -1;
    return WIFEXITED(status) ? WEXITSTATUS(status) : -1;
}

}  // namespace

struct symcpu_program {
    std::vector<uint64_t> argStart, argLength, outStart, outLength;
    std::vector<double> x, result, c;
    std::vector<uint32_t> p, o;
    void* lib = nullptr;
    RunFn run = nullptr;
};

extern "C" {

const char* symcpu_last_error(void) { return g_cpuError.c_str(); }

int symcpu_create(symcpu_t* out, const symhip_program_desc* d) {
    if (!out || !d) return fail(SYMHIP_E_INVALID, "symcpu_create: null argument");
    *out = nullptr;
    auto* h = new symcpu_program();
    h->argStart.assign(d->arg_start, d->arg_start + d->num_args);
    h->argLength.assign(d->arg_length, d->arg_length + d->num_args);
    h->outStart.assign(d->out_start, d->out_start + d->num_outputs);
    h->outLength.assign(d->out_length, d->out_length + d->num_outputs);
    for (uint32_t i = 0; i < d->num_args; ++i)
        if (h->argStart[i] + h->argLength[i] > d->num_values) {
            delete h;
            return fail(SYMHIP_E_INVALID, "symcpu_create: input container outside the value buffer");
        }
    uint64_t numResults = 0;
    for (uint32_t k = 0; k < d->num_outputs; ++k) numResults = std::max<uint64_t>(numResults, h->outStart[k] + h->outLength[k]);
    h->x.assign(d->num_values, 0.0);
    h->result.assign(numResults, 0.0);
    if (d->num_const) h->c.assign(d->const_data, d->const_data + d->num_const);
    if (d->num_index) h->p.assign(d->index_data, d->index_data + d->num_index);
    if (d->num_outpos) h->o.assign(d->outpos_data, d->outpos_data + d->num_outpos);
    *out = h;
    return SYMHIP_SUCCESS;
}

int symcpu_compile(symcpu_t h, const char* source, const char* extra) {
    if (!h || !source) return fail(SYMHIP_E_INVALID, "symcpu_compile: null argument");
    const std::string cxx = compiler();
    std::string flags = kReferenceFlags;
    if (extra) flags += std::string(" ") + extra;
    const std::string key = std::string(source) + "\n//cxx:" + cxx + "\n//flags:" + flags;
    char hash[32];
    std::snprintf(hash, sizeof(hash), "%016llx", (unsigned long long)Sym::CRC64_hash(key.data(), key.size()));
    const std::string dir = cacheDirectory();
    mkdir(dir.c_str(), 0755);
    const std::string base = dir + "/" + hash + "_" + std::to_string(key.size()) + ".cpu";
    const std::string so = base + ".so";
    struct stat st;
    if (stat(so.c_str(), &st) != 0 || st.st_size == 0) {
        const std::string src = base + "." + std::to_string(getpid()) + ".cpp";
        {
            std::ofstream f(src);
            if (!f) return fail(SYMHIP_E_IO, "symcpu_compile: cannot write " + src);
            f << source;
        }
        const std::string tmp = base + "." + std::to_string(getpid()) + ".so";
        std::vector<std::string> argv = {cxx};
        size_t a = 0;
        while (a < flags.size()) {
            const size_t b = flags.find(' ', a);
            const std::string tok = flags.substr(a, b == std::string::npos ? std::string::npos : b - a);
            if (!tok.empty()) argv.push_back(tok);
            if (b == std::string::npos) break;
            a = b + 1;
        }
        const std::string libdir = cxx.substr(0, cxx.rfind("/bin/")) + "/lib";
        for (const char* s : {"-shared", "-fPIC", "-o"}) argv.push_back(s);
        argv.push_back(tmp);
        argv.push_back(src);
        argv.push_back("-Wl,-rpath," + libdir);
        std::string log;
        const int rc = runProcess(argv, log);
        std::remove(src.c_str());
        if (rc != 0) {
            std::remove(tmp.c_str());
            return fail(SYMHIP_E_COMPILE, "symcpu_compile: " + cxx + " failed (" + std::to_string(rc) + "):\n" + log);
        }
        if (std::rename(tmp.c_str(), so.c_str()) != 0) return fail(SYMHIP_E_IO, "symcpu_compile: cannot write " + so);
    }
    if (h->lib) dlclose(h->lib);
    h->lib = dlopen(so.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h->lib) return fail(SYMHIP_E_IO, std::string("symcpu_compile: dlopen: ") + dlerror());
    h->run = (RunFn)dlsym(h->lib, "symlib_cpu_run");
    if (!h->run) return fail(SYMHIP_E_SYMBOL, "symcpu_compile: the library has no symlib_cpu_run");
    return SYMHIP_SUCCESS;
}

int symcpu_set_arg(symcpu_t h, int i, const double* host, size_t n) {
    if (!h || i < 0 || (size_t)i >= h->argStart.size()) return fail(SYMHIP_E_INVALID, "symcpu_set_arg: bad argument index");
    if (n != h->argLength[i]) return fail(SYMHIP_E_INVALID, "symcpu_set_arg: length does not match the input container");
    if (n && !host) return fail(SYMHIP_E_INVALID, "symcpu_set_arg: null data");
    std::memcpy(h->x.data() + h->argStart[i], host, sizeof(double) * n);
    return SYMHIP_SUCCESS;
}

int symcpu_run(symcpu_t h, int threads) {
    if (!h) return fail(SYMHIP_E_INVALID, "symcpu_run: null handle");
    if (!h->run) return fail(SYMHIP_E_NOT_COMPILED, "symcpu_run: program not compiled");
    if (threads < 1) return fail(SYMHIP_E_INVALID, "symcpu_run: threads must be at least 1");
    h->run(h->x.data(), h->x.data(), h->c.data(), h->p.data(), h->o.data(), h->result.data(), threads);
    return SYMHIP_SUCCESS;
}

int symcpu_get_result(symcpu_t h, int k, double* host, size_t n) {
    if (!h || k < 0 || (size_t)k >= h->outStart.size()) return fail(SYMHIP_E_INVALID, "symcpu_get_result: bad output index");
    if (n != h->outLength[k]) return fail(SYMHIP_E_INVALID, "symcpu_get_result: length does not match the output container");
    if (n && !host) return fail(SYMHIP_E_INVALID, "symcpu_get_result: null data");
    std::memcpy(host, h->result.data() + h->outStart[k], sizeof(double) * n);
    return SYMHIP_SUCCESS;
}

void symcpu_destroy(symcpu_t h) {
    if (!h) return;
    if (h->lib) dlclose(h->lib);
    delete h;
}

}  // extern "C"
