#include "Traverse.hpp"

namespace Sym {

// All distinct nodes with operation `op`, in pre-order (src/Traverse.cpp:68-77).
std::vector<Symbolic> harvestOp(OpType op, const Symbolic& x) {
    std::vector<Symbolic> ret;
    preOrderTraverse(
        x,
        [&](const Symbolic& y) {
            if (y.op() == op) ret.push_back(y);
            return true;
        },
        true);
    return ret;
}

}  // namespace Sym
