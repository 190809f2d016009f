// HIP source generation for a compiled Program (new emitter; the reference's HIP text is
// produced by src/ComputeUnit.cpp:368-450 with src/CodeGenerator.cpp:12-118).
//
// Each wave of independent kernels becomes one extern "C" __global__ function.  Its
// workgroups are split into contiguous ranges, one per kernel; a work-item evaluates one
// instance of its kernel's template in registers:
//   gathered inputs   x[p[base + i] + offset]   (structure-of-arrays index columns)
//   constants         c[base + i]
//   local temporaries const double rK = ...;
//   user outputs      out[o[base + i]] = ...    (final position, no gather pass)
//   intermediates     y[base + i] = ...         (read back by a later wave through p)
// All arithmetic is FP64.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <stdexcept>

#include "Program.hpp"
#include "Tiling.hpp"
#include "Traverse.hpp"

namespace Sym {

namespace {

std::string literal(double d) {
    if (std::isnan(d)) return "__builtin_nan(\"\")";
    if (std::isinf(d)) return d > 0 ? "__builtin_huge_val()" : "(-__builtin_huge_val())";
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", d);
    std::string s = buf;
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    if (d < 0 || (d == 0 && std::signbit(d))) s = "(" + s + ")";
    return s;
}

std::string idx(size_t base, bool wide) {
    if (wide) return "(size_t)" + std::to_string(base) + "ull + i";
    return std::to_string(base) + "u + i";
}

struct KernelEmitter {
    const Program& prog;
    const ProgramKernel& K;
    bool wide;
    // tiled form: per-tile table D (operands | ids | outpos), cnt instances, LDS base L
    const TiledProgram* tiled = nullptr;
    size_t kernelIndex = 0;

    std::string tiledLeaf(const std::array<int, 2>& v) const {
        const int slot = v[0];
        const size_t V = K.numVarSlots;
        switch (v[1]) {
            case -2: {
                if (K.userOutputs) {
                    const size_t row = V + (K.numConstSlots ? 1 : 0) + (size_t)slot;
                    return "out[D[" + std::to_string(row) + "u * cnt + j]]";
                }
                return "lds[L + " + std::to_string(slot) + "u * cnt + j]";
            }
            case -3: {
                const auto& va = K.access.at((size_t)slot);
                const int kind = tiled->colKind[tiled->colOffset[kernelIndex] + va.baseVariableId];
                const std::string e = "D[" + std::to_string(va.baseVariableId) + "u * cnt + j]";
                const std::string off = va.offset ? ((va.offset > 0 ? " + " : " - ") + std::to_string(std::abs(va.offset)) + "u") : "";
                if (kind == 0) return "x[" + e + off + "]";
                if (kind == 1) return "lds[" + e + " & 0x7fffffffu]";
                return "gather(x, lds, " + e + ", " + std::to_string(va.offset) + ")";
            }
            case -4:
                return "c[" + std::to_string(K.constOffset + (size_t)slot * K.n) + "u + D[" + std::to_string(K.numVarSlots) +
                       "u * cnt + j]]";
            case LOCAL_VAR:
                return "r" + std::to_string(slot);
            default:
                throw std::logic_error("emitHipTiled: unexpected variable group " + std::to_string(v[1]));
        }
    }

    std::string leaf(const std::array<int, 2>& v) const {
        if (tiled) return tiledLeaf(v);
        const int slot = v[0];
        switch (v[1]) {
            case -2: {
                if (K.userOutputs) return "out[o[" + idx(K.outposOffset + (size_t)slot * K.n, wide) + "]]";
                return "y[" + idx(prog.numInputs + K.interBase + (size_t)slot * K.n, wide) + "]";
            }
            case -3: {
                const auto& va = K.access.at((size_t)slot);
                std::string s = "x[p[" + idx(K.indexOffset + (size_t)va.baseVariableId * K.n, wide) + "]";
                if (va.offset) s += (va.offset > 0 ? " + " : " - ") + std::to_string(std::abs(va.offset)) + (wide ? "" : "u");
                return s + "]";
            }
            case -4:
                return "c[" + idx(K.constOffset + (size_t)slot * K.n, wide) + "]";
            case LOCAL_VAR:
                return "r" + std::to_string(slot);
            default:
                throw std::logic_error("emitHip: unexpected variable group " + std::to_string(v[1]));
        }
    }

    // expression text, built bottom-up without recursion
    std::string expr(const Symbolic& root) const {
        std::vector<std::string> st;
        postOrderTraverse(root, [&](const Symbolic& y) {
            const unsigned nc = y.numChilds();
            const OpType op = y.op();
            if (op == CONST) {
                st.push_back(literal(y.constant()));
                return;
            }
            if (op == VAR) {
                st.push_back(leaf(y.variable()));
                return;
            }
            std::vector<std::string> ch(st.end() - nc, st.end());
            st.erase(st.end() - nc, st.end());
            std::string s;
            auto join = [&](const char* sep) {
                s = "(";
                for (unsigned i = 0; i < nc; ++i) s += (i ? sep : "") + ch[i];
                s += ")";
            };
            switch (op) {
                case ADD:
                    join(" + ");
                    break;
                case MUL:
                case MULC:
                    join(" * ");
                    break;
                case DIV:
                    s = "(" + ch[0] + " / " + ch[1] + ")";
                    break;
                case NEG:
                    s = "(-" + ch[0] + ")";
                    break;
                case RECIPROCAL:
                    s = "(1.0 / " + ch[0] + ")";
                    break;
                case SQRT:
                    s = "sqrt(" + ch[0] + ")";
                    break;
                case POW:
                    s = "pow(" + ch[0] + ", " + ch[1] + ")";
                    break;
                case SIN:
                case COS:
                case TAN:
                case ASIN:
                case ACOS:
                case ATAN:
                case EXP:
                case LOG:
                    s = OpInfos[op].opSymbol + "(" + ch[0] + ")";
                    break;
                case COT:
                    s = "(1.0 / tan(" + ch[0] + "))";
                    break;
                default:
                    throw std::logic_error("emitHip: operation " + OpInfos[op].code + " cannot be emitted");
            }
            st.push_back(std::move(s));
        });
        return st.at(0);
    }

    void statements(const Symbolic& b, std::ostringstream& os, const std::string& ind) const {
        switch (b.op()) {
            case BLOCK:
                for (auto& c : b) statements(c, os, ind);
                break;
            case DEFINE:
                os << ind << "const double r" << b[0].variable()[0] << " = " << expr(b[1]) << ";\n";
                break;
            case ASSIGN: {
                if (b[0].op() != VAR) throw std::logic_error("emitHip: assignment to a non-variable");
                auto v = b[0].variable();
                if (v[1] == -2 && v[0] < 0) break;
                os << ind << leaf(v) << " = " << expr(b[1]) << ";\n";
                break;
            }
            default:
                throw std::logic_error("emitHip: unexpected statement " + OpInfos[b.op()].code);
        }
    }
};

}  // namespace

std::string emitHip(const Program& prog, const EmitOptions& opt, std::vector<LaunchDesc>& launches) {
    const int T = opt.threadsPerBlock;
    if (T <= 0 || T > 1024 || T % 64) throw std::invalid_argument("emitHip: threads per block must be a multiple of 64 up to 1024");
    const size_t maxIdx = std::max({prog.index.size(), prog.consts.size(), prog.outpos.size(), prog.numValues()});
    const bool wide = maxIdx >= (size_t(1) << 31);

    std::ostringstream os;
    os << "// generated by symboliclib_amd: one function per wave of independent kernels, FP64\n";
    os << "typedef unsigned int INDEX;\n\n";

    launches.clear();
    for (int w = 0; w < prog.numWaves; ++w) {
        const std::string name = "symlib_wave" + std::to_string(w);
        os << "extern \"C\" __global__ void __launch_bounds__(" << T << ") " << name
           << "(const double* __restrict__ x, double* __restrict__ y, const double* __restrict__ c,\n"
           << "    const INDEX* __restrict__ p, const INDEX* __restrict__ o, double* __restrict__ out,\n"
           << "    const INDEX* __restrict__ tiles) {\n";
        os << "    const unsigned b = blockIdx.x;\n";
        size_t blocks = 0;
        for (const auto& K : prog.kernels) {
            if (K.wave != w || K.n == 0) continue;
            const size_t nb = (K.n + T - 1) / T;
            const size_t b0 = blocks;
            blocks += nb;
            os << "    if (b < " << blocks << "u) {\n";
            os << "        const unsigned i = (b - " << b0 << "u) * " << T << "u + threadIdx.x;\n";
            os << "        if (i < " << K.n << "u) {\n";
            KernelEmitter ke{prog, K, wide};
            ke.statements(K.code, os, "            ");
            os << "        }\n        return;\n    }\n";
        }
        os << "}\n\n";
        if (blocks > 0xffffffffull) throw std::length_error("emitHip: grid too large");
        launches.push_back({name, (uint32_t)blocks, (uint32_t)T, 0});
    }
    return os.str();
}

std::string emitHipTiled(const Program& prog, const TiledProgram& tp, const EmitOptions& opt,
                         std::vector<LaunchDesc>& launches) {
    if (!tp.valid) throw std::invalid_argument("emitHipTiled: program is not tiled");
    const int T = opt.threadsPerBlock;
    std::ostringstream os;
    os << "// generated by symboliclib_amd: tile-fused program, one workgroup per tile, FP64\n";
    os << "typedef unsigned int INDEX;\n\n";
    os << "__device__ __forceinline__ double gather(const double* __restrict__ x, const double* lds, INDEX v, int off) {\n"
       << "    return (v & 0x80000000u) ? lds[v & 0x7fffffffu] : x[(int)v + off];\n}\n\n";
    os << "extern \"C\" __global__ void __launch_bounds__(" << T << ") symlib_tiled"
       << "(const double* __restrict__ x, double* __restrict__ y, const double* __restrict__ c,\n"
       << "    const INDEX* __restrict__ p, const INDEX* __restrict__ o, double* __restrict__ out,\n"
       << "    const INDEX* __restrict__ tiles) {\n";
    os << "    extern __shared__ double lds[];\n";
    os << "    const INDEX* __restrict__ H = tiles + tiles[blockIdx.x];\n";
    for (int w = 0; w < prog.numWaves; ++w) {
        std::vector<size_t> ks;
        for (size_t r = 0; r < prog.kernels.size(); ++r)
            if (prog.kernels[r].wave == w && tp.kernelUsed[r] && prog.kernels[r].n) ks.push_back(r);
        if (ks.empty()) continue;
        if (w > 0) os << "    __syncthreads();\n";
        os << "    {\n";
        std::string total;
        for (size_t q = 0; q < ks.size(); ++q) {
            os << "        const unsigned n" << q << " = H[" << 3 * ks[q] << "];\n";
            total += (q ? " + n" : "n") + std::to_string(q);
        }
        os << "        const unsigned total = " << total << ";\n";
        os << "        for (unsigned g = threadIdx.x; g < total; g += " << T << "u) {\n";
        os << "            unsigned j = g;\n";
        for (size_t q = 0; q < ks.size(); ++q) {
            const size_t r = ks[q];
            os << "            if (j < n" << q << ") {\n";
            os << "                const unsigned cnt = n" << q << ";\n";
            os << "                const INDEX* __restrict__ D = tiles + H[" << 3 * r + 1 << "];\n";
            if (!prog.kernels[r].userOutputs) os << "                const unsigned L = H[" << 3 * r + 2 << "];\n";
            KernelEmitter ke{prog, prog.kernels[r], false, &tp, r};
            ke.statements(prog.kernels[r].code, os, "                ");
            os << "                continue;\n            }\n";
            os << "            j -= n" << q << ";\n";
        }
        os << "        }\n    }\n";
    }
    os << "}\n";
    launches.clear();
    launches.push_back({"symlib_tiled", (uint32_t)tp.numTiles, (uint32_t)T, (uint32_t)tp.ldsBytes});
    return os.str();
}

}  // namespace Sym
