// Driver around the reference Integer_Heap (include/integer_heap.cuh), built by
// oracle/Makefile into oracle/_ref/heap_trace.  Reads a scripted trace on stdin:
//   n_nodes buckets n_ops, then n_ops lines "code node val" (0 push, 1 modify,
//   2 pop, 3 refresh, 4 PMAX = val)
// and prints the popped node ids, one per line.  Used only to write
// tests/golden/heap_traces.json.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>
#include "integer_heap.cuh"

int main()
{
    int n_nodes = 0, buckets = 0, n_ops = 0;
    if (std::scanf("%d %d %d", &n_nodes, &buckets, &n_ops) != 3) return 1;
    std::vector<std::unique_ptr<Node>> nodes(n_nodes);
    for (int i = 0; i < n_nodes; ++i) nodes[i] = std::make_unique<Node>(i);
    Integer_Heap heap(buckets);
    for (int i = 0; i < n_ops; ++i) {
        int code, node, val;
        if (std::scanf("%d %d %d", &code, &node, &val) != 3) return 2;
        if (code == 0) { Node* p = nodes[node].get(); heap.push(p); }
        else if (code == 1) { Node* p = nodes[node].get(); heap.modify(p, val); }
        else if (code == 2) { Node* p = nullptr; heap.pop(p); std::printf("%d\n", p ? p->ID : -1); }
        else if (code == 3) { Node* p = nodes[node].get(); heap.refresh(p); }
        else heap.PMAX = val;
    }
    return 0;
}
