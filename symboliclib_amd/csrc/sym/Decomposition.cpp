#include "Decomposition.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <unordered_map>

#include "IdMap.hpp"
#include "Template.hpp"
#include "Traverse.hpp"
#include "Utilities.hpp"

namespace Sym {

namespace {

constexpr unsigned char MERGED = 1;
constexpr unsigned char EXPLICITLY_MERGED = 2;
constexpr unsigned char FIXED_NODE = 3;

// Children of a node keyed by algebraic hash, kept sorted (iteration in key order, as an
// ordered map); nodes have few children, so a flat vector beats a tree.
template <class T>
class FlatMap {
    std::vector<std::pair<hash_t, T>> v_;

public:
    using iterator = typename std::vector<std::pair<hash_t, T>>::iterator;
    iterator begin() { return v_.begin(); }
    iterator end() { return v_.end(); }
    iterator find(hash_t k) {
        auto it = std::lower_bound(v_.begin(), v_.end(), k, [](const auto& e, hash_t key) { return e.first < key; });
        return (it != v_.end() && it->first == k) ? it : v_.end();
    }
    bool insert(const std::pair<hash_t, T>& e) {
        auto it = std::lower_bound(v_.begin(), v_.end(), e.first, [](const auto& x, hash_t key) { return x.first < key; });
        if (it != v_.end() && it->first == e.first) return false;
        v_.insert(it, e);
        return true;
    }
    void erase(hash_t k) {
        auto it = find(k);
        if (it != v_.end()) v_.erase(it);
    }
    size_t size() const { return v_.size(); }
};

// Small identity cache: linear scan while short, hash map beyond.
class ExtractCache {
    std::vector<std::pair<long long, std::pair<Symbolic, hash_t>>> small_;
    std::unordered_map<long long, std::pair<Symbolic, hash_t>, IdentityHash<long long>> big_;
    bool useBig_ = false;

public:
    const std::pair<Symbolic, hash_t>* find(long long k) const {
        if (useBig_) {
            auto it = big_.find(k);
            return it == big_.end() ? nullptr : &it->second;
        }
        for (auto& e : small_)
            if (e.first == k) return &e.second;
        return nullptr;
    }
    void put(long long k, const std::pair<Symbolic, hash_t>& v) {
        if (!useBig_ && small_.size() < 32) {
            small_.emplace_back(k, v);
            return;
        }
        if (!useBig_) {
            for (auto& e : small_) big_.emplace(e.first, e.second);
            small_.clear();
            useBig_ = true;
        }
        big_[k] = v;
    }
};

// Rebuild node `y` with new children (keeps the template id of INSTANCE nodes).
Symbolic rebuild(const Symbolic& y, const std::vector<Symbolic>& childs) {
    if (y.op() == INSTANCE) return Symbolic::instance(y.tag(), childs, 0);
    return Symbolic(y.op(), childs);
}

class ExpressionNode {
public:
    Symbolic x;
    int id;
    Symbolic fullx;
    unsigned char status = 0;
    unsigned int level = (unsigned)-1;
    unsigned char complexity = 0;
    hash_t structureHash = 0;
    unsigned int numParents = 0;
    FlatMap<ExpressionNode*> childs;

    ExpressionNode(const Symbolic& fx, int id_) : id(id_), fullx(fx) {}

    hash_t hash() const { return fullx.algebraicHash(); }

    // ids of all non-merged descendants reached through merged nodes
    std::vector<int> allChildren() {
        std::vector<int> ret;
        std::vector<ExpressionNode*> stack{this};
        while (!stack.empty()) {
            auto curr = stack.back();
            stack.pop_back();
            for (auto& c : curr->childs) {
                if (c.second->status == MERGED) stack.push_back(c.second);
                else ret.push_back(c.second->id);
            }
        }
        deleteDuplicates(ret);
        return ret;
    }

    void addChild(const std::pair<hash_t, ExpressionNode*>& c) {
        if (childs.insert(c)) ++c.second->numParents;
    }

    // absorb child n into this node's computation
    void merge(ExpressionNode& n) {
        childs.erase(n.hash());
        if (n.status != EXPLICITLY_MERGED) {
            for (auto& c : n.childs) --c.second->numParents;
            n.numParents = 0;
            n.status = EXPLICITLY_MERGED;
        }
        for (auto& c : n.childs) addChild(c);  // n is never this node
        complexity = (unsigned char)std::min(255, complexity + n.complexity);
    }

    Symbolic extractExpression(const int variableGroupId) {
        if (x.op() != NOOP) return x;

        std::vector<std::pair<Symbolic, hash_t>> stack;
        ExtractCache cache;
        std::vector<hash_t> hashes;

        auto preFun = [&](const Symbolic& y) {
            if (auto* hit = cache.find(y.id())) {
                stack.push_back(*hit);
                return false;
            }
            if (y.op() == FIXED) {
                stack.push_back(std::make_pair(y[0], VarHash));
                return false;
            }
            auto it = childs.find(y.ahash());
            if (it != childs.end()) {
                if (it->second->status == MERGED) {
                    const Symbolic xpr = it->second->extractExpression(variableGroupId);
                    stack.push_back(std::make_pair(xpr, it->second->structureHash));
                } else {
                    stack.push_back(std::make_pair(Symbolic(it->second->id, variableGroupId), VarHash));
                }
                return false;
            }
            return true;
        };

        auto postFun = [&](const Symbolic& y) {
            const unsigned nc = y.numChilds();
            if (nc == 0) {
                stack.push_back(std::make_pair(y, y.op() == VAR ? VarHash : ConstHash));
                return;
            }
            if (OpInfos[y.op()].commutative)
                sortBy(stack.end() - nc, stack.end(), [](const auto& e) { return e.second; });

            hashes.clear();
            const size_t first = stack.size() - nc;
            for (size_t i = first; i < stack.size(); ++i) hashes.push_back(stack[i].second);
            hashes.push_back(OpInfos[y.op()].hash);
            if (y.op() == INSTANCE) hashes.push_back((hash_t)y.tag());
            static_assert(sizeof(std::pair<Symbolic, hash_t>) == 2 * sizeof(void*), "layout");
            // children are the first members of consecutive stack entries
            std::vector<Symbolic> ch;
            ch.reserve(nc);
            for (size_t i = first; i < stack.size(); ++i) ch.push_back(std::move(stack[i].first));
            stack.erase(stack.end() - nc, stack.end());
            stack.push_back(std::make_pair(rebuild(y, ch), Sym::hash(hashes)));
            cache.put(y.id(), stack.back());
        };

        prePostOrderTraverse(fullx, preFun, postFun);
        assert(stack.size() == 1);
        x = stack.front().first;
        structureHash = stack.front().second;

        if (x.op() != BLOCK && x.op() != ASSIGN && x.op() != INSTANCE && status != MERGED)
            x = Symbolic(ASSIGN, Symbolic(id, variableGroupId), x);
        return x;
    }
};

}  // namespace

std::vector<ExpressionBlock> decompose(const std::vector<Symbolic>& expr, const unsigned char complexityThreshold,
                                       const unsigned char variableGroupId, const bool mergeSingleParent) {
    if (complexityThreshold == 255) {
        std::vector<ExpressionBlock> ret;
        ret.reserve(expr.size());
        for (auto& e : expr) ret.emplace_back(e);
        return ret;
    }

    std::vector<ExpressionNode> allNodes;
    allNodes.reserve(2 * expr.size() + 16);
    // node identity -> (occurrence count, index + 1 in allNodes)
    IdMap<std::pair<unsigned, unsigned>> exprCount(4 * expr.size() + 16);
    int id = 0;
    const bool timing = std::getenv("SYMLIB_TIMING") != nullptr && expr.size() > 100000;
    auto tprev = std::chrono::steady_clock::now();
    auto phase = [&](const char* what) {
        if (!timing) return;
        auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[symlib]   decompose: %-18s %8.3f s\n", what, std::chrono::duration<double>(now - tprev).count());
        tprev = now;
    };

    // 1. count occurrences of every sufficiently complex sub-expression
    for (const auto& x : expr) {
        preOrderTraverse(x, [&](const Symbolic& y) {
            if (y.op() == FIXED) {
                auto& xc = exprCount[y[1].id()];
                if (xc.first++ == 0) {
                    allNodes.emplace_back(y[1], id++);
                    allNodes.back().status = FIXED_NODE;
                    xc.second = (unsigned)allNodes.size();
                }
                return false;
            } else if (y.complexity() < complexityThreshold) {
                return false;
            } else {
                auto& xc = exprCount[y.id()];
                if (xc.first == 1) {
                    allNodes.emplace_back(y, id++);
                    xc.second = (unsigned)allNodes.size();
                }
                return xc.first++ == 0;
            }
        });
    }

    phase("count");
    // 2. every output is a node
    std::vector<unsigned> outputNodes;
    outputNodes.reserve(expr.size());
    for (auto& x : expr) {
        allNodes.emplace_back(x, id++);
        allNodes.back().numParents++;
        outputNodes.push_back((unsigned)allNodes.size() - 1);
        exprCount[x.id()].second = (unsigned)allNodes.size();
    }

    phase("outputs");
    // 3. link nodes to the nodes they contain and estimate their own cost
    for (auto& node : allNodes) {
        if (node.status == FIXED_NODE) {
            node.complexity = 255;
            continue;
        }
        int comp = 0;
        preOrderTraverse(node.fullx, [&](const Symbolic& y) {
            if (y.op() == FIXED) {
                auto& data = exprCount[y[1].id()];
                node.addChild(std::make_pair(y[1].ahash(), allNodes.data() + data.second - 1));
                return false;
            }
            if (y.complexity() < complexityThreshold) {
                comp += y.complexity();
                return false;
            }
            auto* it = exprCount.find(y.id());
            const unsigned second = it ? it->second : 0;
            if (second != 0 && y.ahash() != node.hash()) {
                node.addChild(std::make_pair(y.ahash(), allNodes.data() + second - 1));
                return false;
            }
            comp += std::max(1, (int)y.numChilds()) * OpInfos[y.op()].cost;
            return true;
        });
        node.complexity = (unsigned char)std::min(255, comp);
    }

    phase("children");
    // 4. merge cheap or single-parent children into their parents and assign levels
    std::vector<ExpressionNode*> stack;
    // membership flags by node index; iteration in index order = allNodes order
    std::vector<char> reachable(allNodes.size(), 0);
    auto reach = [&](ExpressionNode* n) {
        char& r = reachable[n - allNodes.data()];
        const bool fresh = !r;
        r = 1;
        return fresh;
    };
    std::vector<ExpressionNode*> childStack;

    for (auto i : outputNodes) {
        ExpressionNode* x = allNodes.data() + i;
        stack.clear();
        stack.push_back(x);
        x->level = 0;

        while (!stack.empty()) {
            ExpressionNode* node = stack.back();
            stack.pop_back();

            if (node->status == FIXED_NODE) {
                reach(node);
            } else if (node->status != 0 || reach(node)) {
                childStack.clear();
                for (auto& c : node->childs) childStack.push_back(c.second);

                while (!childStack.empty()) {
                    ExpressionNode* cnode = childStack.back();
                    childStack.pop_back();
                    if (cnode->status == FIXED_NODE) {
                    } else if (cnode->status == EXPLICITLY_MERGED || cnode->complexity < complexityThreshold) {
                        node->merge(*cnode);
                        for (auto& c : cnode->childs) childStack.push_back(c.second);
                    } else if (mergeSingleParent && cnode->numParents == 1) {
                        cnode->status = MERGED;
                        for (auto& c : cnode->childs) stack.push_back(c.second);
                    }
                }

                for (auto& c : node->childs) {
                    c.second->level = node->level + 1;
                    stack.push_back(c.second);
                }
            }
        }
    }

    phase("merge/levels");
    // 5. extract one block per remaining node
    std::vector<int> blockIdMap(allNodes.size(), -1);
    std::vector<ExpressionBlock> blocks;
    for (size_t xi = 0; xi < allNodes.size(); ++xi) {
        if (!reachable[xi]) continue;
        ExpressionNode* x = allNodes.data() + xi;
        if (x->status != MERGED) {
            blockIdMap[x->id] = (int)blocks.size();
            ExpressionBlock block;
            block.level = (int)x->level;
            block.x = x->extractExpression(variableGroupId);
            block.structureHash = x->structureHash;
            block.childs = x->allChildren();
            blocks.push_back(std::move(block));
        }
    }

    phase("extract");
    for (auto& b : blocks)
        for (auto& i : b.childs) i = blockIdMap[i];

    return blocks;
}

}  // namespace Sym
