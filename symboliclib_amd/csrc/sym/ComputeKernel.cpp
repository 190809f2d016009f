#include "ComputeKernel.hpp"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>

#include "Decomposition.hpp"
#include "Simplify.hpp"
#include "Template.hpp"

namespace Sym {

namespace {

using Column = std::vector<Symbolic>;  // the nodes of every instance at one template position

// Builds the template of a group by walking its instances in lock step.
class GroupWalker {
public:
    GroupWalker(const std::vector<Symbolic>& instances, KernelTemplate& K) : roots_(instances), K_(K) {}

    Symbolic run() {
        struct Frame {
            Column col;
            bool outputPosition;
            unsigned next = 0;
            std::vector<Symbolic> childs;
        };
        std::vector<Frame> frames;
        frames.push_back({roots_, false, 0, {}});
        Symbolic result;
        while (!frames.empty()) {
            Frame& f = frames.back();
            if (f.next == 0 && f.childs.empty()) {
                checkUniform(f.col);
                const hash_t sig = signature(f.col);
                auto hit = shared_.find(sig);
                Symbolic leaf;
                if (hit != shared_.end()) leaf = hit->second;
                else if (f.col.front().op() == VAR) leaf = variableLeaf(f.col, f.outputPosition);
                else if (f.col.front().op() == CONST) leaf = constantLeaf(f.col);
                if (leaf.op() != NOOP) {
                    if (hit == shared_.end()) shared_.emplace(sig, leaf);
                    if (!deliver(frames, leaf, result)) break;
                    continue;
                }
            }
            const Symbolic& head = f.col.front();
            if (f.next < head.numChilds()) {
                const unsigned k = f.next++;
                // output positions: the left side of an assignment, the output arguments of an
                // expression template
                bool out = head.op() == ASSIGN && k == 0;
                if (head.op() == INSTANCE) out = (int)k < getTemplate(head.tag()).numOutputs;
                Column child;
                child.reserve(f.col.size());
                for (const Symbolic& x : f.col) child.push_back(x[k]);
                frames.push_back({std::move(child), out, 0, {}});
                continue;
            }
            const Symbolic node =
                head.op() == INSTANCE ? Symbolic::instance(head.tag(), f.childs, 0) : Symbolic(head.op(), f.childs);
            shared_.emplace(signature(f.col), node);
            if (!deliver(frames, node, result)) break;
        }
        return result;
    }

private:
    const std::vector<Symbolic>& roots_;
    KernelTemplate& K_;
    std::unordered_map<hash_t, Symbolic> shared_;  // column signature -> template node

    // hand a finished node to the frame below; false when it was the root
    template <class Frames>
    static bool deliver(Frames& frames, const Symbolic& node, Symbolic& result) {
        frames.pop_back();
        if (frames.empty()) {
            result = node;
            return false;
        }
        frames.back().childs.push_back(node);
        return true;
    }

    static hash_t signature(const Column& col) {
        std::vector<hash_t> h;
        h.reserve(col.size() + 1);
        for (const Symbolic& x : col) h.push_back(x.ahash());
        h.push_back((hash_t)col.size());
        return hash(h);
    }

    static void checkUniform(const Column& col) {
        const Symbolic& a = col.front();
        for (const Symbolic& b : col)
            if (b.op() != a.op() || b.numChilds() != a.numChilds() || (a.op() == INSTANCE && b.tag() != a.tag()))
                throw std::logic_error("kernel: the instances of a group differ in structure");
    }

    Symbolic variableLeaf(const Column& col, bool output) {
        std::vector<std::array<int, 2>> vars;
        vars.reserve(col.size());
        for (const Symbolic& x : col) vars.push_back(x.variable());
        if (output) {
            K_.outputColumns.push_back(std::move(vars));
            return Symbolic((int)K_.outputColumns.size() - 1, -2);
        }
        pendingInputs_.push_back(std::move(vars));
        return Symbolic((int)pendingInputs_.size() - 1, -3);
    }

    Symbolic constantLeaf(const Column& col) {
        const double c0 = col.front().constant();
        bool uniform = true;
        std::vector<double> vals;
        vals.reserve(col.size());
        for (const Symbolic& x : col) {
            vals.push_back(x.constant());
            uniform = uniform && x.constant() == c0;
        }
        if (uniform) return Symbolic(c0);
        K_.constantColumns.push_back(std::move(vals));
        return Symbolic((int)K_.constantColumns.size() - 1, -4);
    }

public:
    // input leaves in template order; stored after compression
    std::vector<std::vector<std::array<int, 2>>> pendingInputs_;
};

// Store the input columns, each column that is another stored column shifted by a constant
// (one input group in every instance) as a reference to it.
void storeInputColumns(std::vector<std::vector<std::array<int, 2>>>& leaves, KernelTemplate& K) {
    std::unordered_map<hash_t, std::vector<int>> byPattern;  // (group, offsets from the first) -> stored columns
    for (auto& col : leaves) {
        const int g = col.front()[1];
        bool oneGroup = g != GLOBAL_INTERMEDIATE && g != EXPLICIT_INTERMEDIATE;
        std::vector<long long> pattern{g};
        for (const auto& v : col) {
            oneGroup = oneGroup && v[1] == g;
            pattern.push_back((long long)v[0] - col.front()[0]);
        }
        int found = -1;
        hash_t key = 0;
        if (oneGroup) {
            key = hash(pattern);
            for (int s : byPattern[key]) {
                const auto& stored = K.inputColumns[(size_t)s];
                bool same = stored.size() == col.size() && stored.front()[1] == g;
                for (size_t i = 1; same && i < col.size(); ++i)
                    same = (long long)stored[i][0] - stored.front()[0] == pattern[i + 1];
                if (same) {
                    found = s;
                    break;
                }
            }
        }
        if (found >= 0) {
            K.inputLeaves.push_back({found, col.front()[0] - K.inputColumns[(size_t)found].front()[0]});
            continue;
        }
        if (oneGroup) byPattern[key].push_back((int)K.inputColumns.size());
        K.inputLeaves.push_back({(int)K.inputColumns.size(), 0});
        K.inputColumns.push_back(std::move(col));
    }
}

}  // namespace

KernelTemplate buildKernelTemplate(const std::vector<Symbolic>& instances, bool simplifyTemplate) {
    if (instances.empty()) throw std::invalid_argument("kernel: empty group");
    KernelTemplate K;
    K.instances = instances.size();
    GroupWalker walker(instances, K);
    const Symbolic t = walker.run();
    storeInputColumns(walker.pendingInputs_, K);
    K.code = optimizeTemplate(t, simplifyTemplate);
    return K;
}

Symbolic optimizeTemplate(Symbolic t, bool simplifyTemplate) {
    if (t.op() == INSTANCE) t = expandInstance(t);
    t = foldConstants(t);
    std::vector<DecomposedBlock> blocks = decompose({t}, 0, LOCAL_VAR, false);
    // a block only reads blocks of lower height: ascending heights put definitions first
    std::stable_sort(blocks.begin(), blocks.end(),
                     [](const DecomposedBlock& a, const DecomposedBlock& b) { return a.height < b.height; });
    std::vector<Symbolic> statements;
    statements.reserve(blocks.size());
    for (const DecomposedBlock& b : blocks) {
        Symbolic s = simplifyTemplate ? simplify(b.code) : b.code;
        if (s.op() == ASSIGN && s[0].op() == VAR && s[0].variable()[1] == LOCAL_VAR) s = Symbolic(DEFINE, s[0], s[1]);
        statements.push_back(s);
    }
    return Symbolic(BLOCK, statements);
}

}  // namespace Sym
