// Decomposition into blocks, in five passes over an explicit graph of candidate nodes:
//   1. census: walk every root, entering each distinct node (by identity) once, and count how
//      often each node of complexity >= threshold is reached (by identity: a value written out
//      twice as two separate trees is computed twice); a FIXED sub-expression names its sealed
//      BLOCK, which is a node of its own and is not entered;
//   2. graph: the nodes are the roots, the sealed blocks and the sub-expressions counted at
//      least twice; walking a node's expression down to the next nodes gives its edges and its
//      own cost (operations outside other nodes);
//   3. inlining: a node that is cheap on its own, or (optionally) read by one node only, is
//      written into its readers instead of becoming a block;
//   4. blocks: the remaining nodes, each reading the blocks reached through inlined nodes;
//      heights by a post-order walk of that graph;
//   5. extraction: each block's expression with nodes that are blocks replaced by their
//      variables, the operands of commutative operations ordered by shape hash, and its shape.
#include "Decomposition.hpp"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>
#include <utility>

namespace Sym {

namespace {

struct Node {
    Symbolic expr;     // the sub-expression (a sealed node: its BLOCK)
    bool root = false;
    bool sealed = false;
    int ownCost = 0;
    std::vector<int> edges;  // nodes reached directly, each once
    int readers = 0;         // distinct nodes (or the caller, for roots) that reach this one
    bool inlined = false;
    int block = -1;
};

class Decomposer {
public:
    Decomposer(unsigned char threshold, int group, bool inlineSingleReader)
        : threshold_(threshold), group_(group), inlineSingle_(inlineSingleReader) {}

    std::vector<DecomposedBlock> run(const std::vector<Symbolic>& roots) {
        census(roots);
        for (const Symbolic& r : roots) {
            Node n;
            n.expr = r;
            n.root = true;
            n.readers = 1;
            nodes_.push_back(n);
        }
        for (size_t i = 0; i < nodes_.size(); ++i) linkNode((int)i);
        for (Node& n : nodes_)
            for (int c : n.edges) ++nodes_[(size_t)c].readers;
        for (Node& n : nodes_)
            if (!n.root && !n.sealed) n.inlined = n.ownCost < threshold_ || (inlineSingle_ && n.readers == 1);
        return buildBlocks();
    }

private:
    const int threshold_;
    const int group_;
    const bool inlineSingle_;
    std::vector<Node> nodes_;
    std::unordered_map<hash_t, int> nodeOf_;  // algebraic hash -> node (shared and sealed nodes)

    bool tracked(const Symbolic& y) const { return y.complexity() >= threshold_; }

    // pass 1: occurrence counts; a node is made at its second occurrence (sealed blocks at the
    // first), so the node order follows the walk; later nodes algebraically equal to one made
    // before are that node
    void census(const std::vector<Symbolic>& roots) {
        std::unordered_map<long long, unsigned> seen;
        std::unordered_set<long long> entered;
        std::vector<Symbolic> todo;
        for (const Symbolic& root : roots) {
            todo.push_back(root);
            while (!todo.empty()) {
                const Symbolic y = todo.back();
                todo.pop_back();
                if (y.op() == FIXED) {
                    const hash_t h = y[1].ahash();
                    if (!nodeOf_.count(h)) {
                        Node n;
                        n.expr = y[1];
                        n.sealed = true;
                        nodeOf_.emplace(h, (int)nodes_.size());
                        nodes_.push_back(n);
                    }
                    continue;
                }
                if (!tracked(y)) continue;
                const hash_t h = y.ahash();
                if (++seen[y.id()] == 2 && !nodeOf_.count(h)) {
                    Node n;
                    n.expr = y;
                    nodeOf_.emplace(h, (int)nodes_.size());
                    nodes_.push_back(n);
                }
                if (!entered.insert(y.id()).second) continue;
                for (unsigned k = y.numChilds(); k-- > 0;) todo.push_back(y[k]);
            }
        }
    }

    // the node a sub-expression stands for, or -1 (never `self`: an expression algebraically
    // equal to its own node, such as 1 * x inside x, is walked through)
    int nodeFor(const Symbolic& y, int self) const {
        if (y.op() == FIXED) return nodeOf_.at(y[1].ahash());
        if (!tracked(y)) return -1;
        auto it = nodeOf_.find(y.ahash());
        return it == nodeOf_.end() || it->second == self ? -1 : it->second;
    }

    // pass 2: edges and own cost of node i
    void linkNode(int i) {
        Node& n = nodes_[(size_t)i];
        if (n.sealed) {
            n.ownCost = 255;
            return;
        }
        std::unordered_set<int> edges;
        int cost = 0;
        std::vector<Symbolic> todo{n.expr};
        while (!todo.empty()) {
            const Symbolic y = todo.back();
            todo.pop_back();
            const int c = nodeFor(y, i);
            if (c >= 0) {
                if (edges.insert(c).second) n.edges.push_back(c);
                continue;
            }
            if (!tracked(y)) {
                cost += y.complexity();
                continue;
            }
            cost += std::max(1, (int)y.numChilds()) * OpInfos[y.op()].cost;
            for (unsigned k = y.numChilds(); k-- > 0;) todo.push_back(y[k]);
        }
        n.ownCost = std::min(cost, 255);
    }

    // pass 4 and 5
    std::vector<DecomposedBlock> buildBlocks() {
        std::vector<int> blockNodes;
        for (size_t i = 0; i < nodes_.size(); ++i)
            if (!nodes_[i].inlined) {
                nodes_[i].block = (int)blockNodes.size();
                blockNodes.push_back((int)i);
            }
        // blocks read by each node, looking through inlined nodes
        std::vector<std::vector<int>> reads(nodes_.size());
        std::vector<char> state(nodes_.size(), 0);  // 0 new, 1 open, 2 done
        for (size_t start = 0; start < nodes_.size(); ++start) {
            if (state[start]) continue;
            std::vector<std::pair<int, size_t>> path{{(int)start, 0}};
            state[start] = 1;
            while (!path.empty()) {
                auto& [v, next] = path.back();
                const Node& n = nodes_[(size_t)v];
                if (next < n.edges.size()) {
                    const int c = n.edges[next++];
                    if (state[(size_t)c] == 1) throw std::logic_error("decompose: cyclic sub-expressions");
                    if (!state[(size_t)c]) {
                        state[(size_t)c] = 1;
                        path.push_back({c, 0});
                    }
                    continue;
                }
                std::vector<int> r;
                for (int c : n.edges) {
                    const Node& cn = nodes_[(size_t)c];
                    if (cn.inlined) r.insert(r.end(), reads[(size_t)c].begin(), reads[(size_t)c].end());
                    else r.push_back(cn.block);
                }
                std::sort(r.begin(), r.end());
                r.erase(std::unique(r.begin(), r.end()), r.end());
                reads[(size_t)v] = std::move(r);
                state[(size_t)v] = 2;
                path.pop_back();
            }
        }
        std::vector<DecomposedBlock> blocks(blockNodes.size());
        for (size_t b = 0; b < blockNodes.size(); ++b) {
            blocks[b].reads = reads[(size_t)blockNodes[b]];
            extract(blockNodes[b], blocks[b]);
        }
        // heights: every block's reads have lower indices in a post-order of the read graph
        std::vector<int> height(blocks.size(), -1);
        for (size_t s = 0; s < blocks.size(); ++s) {
            std::vector<std::pair<int, size_t>> path{{(int)s, 0}};
            while (!path.empty()) {
                auto& [b, next] = path.back();
                if (height[(size_t)b] >= 0) {
                    path.pop_back();
                    continue;
                }
                const auto& r = blocks[(size_t)b].reads;
                if (next < r.size()) {
                    const int c = r[next++];
                    if (height[(size_t)c] < 0) path.push_back({c, 0});
                    continue;
                }
                int h = 0;
                for (int c : r) h = std::max(h, height[(size_t)c] + 1);
                height[(size_t)b] = h;
                path.pop_back();
            }
        }
        for (size_t b = 0; b < blocks.size(); ++b) blocks[b].height = height[b];
        return blocks;
    }

    // pass 5: the expression of block node i and its shape
    void extract(int i, DecomposedBlock& out) const {
        const Node& n = nodes_[(size_t)i];
        struct Built {
            Symbolic x;
            hash_t shape;
        };
        std::unordered_map<long long, Built> done;  // by identity, within this block
        struct Frame {
            Symbolic y;
            unsigned next;
        };
        std::vector<Frame> frames{{n.expr, 0}};
        std::vector<Built> results;
        while (!frames.empty()) {
            Frame& f = frames.back();
            const Symbolic y = f.y;
            if (f.next == 0) {
                // leaves and replacements
                auto hit = done.find(y.id());
                if (hit != done.end()) {
                    results.push_back(hit->second);
                    frames.pop_back();
                    continue;
                }
                if (y.op() == FIXED) {
                    // the sealed block assigns this variable
                    results.push_back({y[0], VarHash});
                    frames.pop_back();
                    continue;
                }
                // a sealed block is taken whole; elsewhere a node that is a block is read
                // through its variable, an inlined one is written out in place
                const int c = n.sealed ? -1 : nodeFor(y, i);
                if (c >= 0 && !nodes_[(size_t)c].inlined) {
                    results.push_back({Symbolic(c, group_), VarHash});
                    frames.pop_back();
                    continue;
                }
                if (y.numChilds() == 0) {
                    results.push_back({y, y.op() == VAR ? VarHash : ConstHash});
                    frames.pop_back();
                    continue;
                }
            }
            if (f.next < y.numChilds()) {
                const Symbolic child = y[f.next++];
                frames.push_back({child, 0});
                continue;
            }
            // all operands built: order commutative operands by shape, then hash the shape
            const unsigned nc = y.numChilds();
            std::vector<Built> ops(results.end() - nc, results.end());
            results.erase(results.end() - nc, results.end());
            if (OpInfos[y.op()].commutative)
                std::stable_sort(ops.begin(), ops.end(), [](const Built& a, const Built& b) { return a.shape < b.shape; });
            std::vector<hash_t> key;
            key.reserve(nc + 2);
            std::vector<Symbolic> childs;
            childs.reserve(nc);
            for (const Built& o : ops) {
                key.push_back(o.shape);
                childs.push_back(o.x);
            }
            key.push_back(OpInfos[y.op()].hash);
            if (y.op() == INSTANCE) key.push_back((hash_t)y.tag());
            const Symbolic rebuilt = y.op() == INSTANCE ? Symbolic::instance(y.tag(), childs, 0) : Symbolic(y.op(), childs);
            const Built b{rebuilt, hash(key)};
            done.emplace(y.id(), b);
            results.push_back(b);
            frames.pop_back();
        }
        Symbolic x = results.front().x;
        out.shape = results.front().shape;
        if (x.op() != BLOCK && x.op() != ASSIGN && x.op() != INSTANCE) x = Symbolic(ASSIGN, Symbolic(i, group_), x);
        out.code = x;
    }
};

}  // namespace

std::vector<DecomposedBlock> decompose(const std::vector<Symbolic>& roots, unsigned char threshold, int group,
                                       bool inlineSingleReader) {
    if (threshold == 255) {
        std::vector<DecomposedBlock> blocks(roots.size());
        for (size_t i = 0; i < roots.size(); ++i) {
            blocks[i].code = roots[i];
            blocks[i].shape = roots[i].computeStructureHash();
        }
        return blocks;
    }
    return Decomposer(threshold, group, inlineSingleReader).run(roots);
}

}  // namespace Sym
