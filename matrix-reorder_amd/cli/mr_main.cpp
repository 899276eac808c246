// mr_main.cpp -- the reference's experiment driver (src/main.cu:1-152) on
// the C++ mirror of its GRAPH interface (include/mr_graph.hpp).
//
//   mr_main <graph id> <type> [width1 width2]
//
// type: origin | degree | dbg | rcm | rabbitorder | gorder1 | gorder5 |
// gorder6 | order (order takes width1 width2, as main.cu:129-133).  The graph
// id is looked up in the config file (lines "id path rows cols nnz"); records
// are appended to the log file in the reference's format (graph id, the
// method's fields, then the MERBIT spmv fields).  The reference hard-codes
// its config, log and device (main.cu:6-8); here they come from MR_CONFIG,
// MR_LOG and MR_DEVICE (defaults config.txt, log.txt, 0).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "mr_graph.hpp"

int main(int argc, char* argv[])
{
    if (argc < 3) {
        std::fprintf(stderr,
                     "usage: %s <graph id> <origin|degree|dbg|rcm|rabbitorder|gorder1|gorder5|gorder6|order> "
                     "[width1 width2]\n",
                     argv[0]);
        return 2;
    }
    const char* cfg = std::getenv("MR_CONFIG");
    const char* lg = std::getenv("MR_LOG");
    const char* dev = std::getenv("MR_DEVICE");
    const int gid = std::stoi(argv[1]);
    const std::string type = argv[2];
    if (type == "order" && argc < 5) {
        std::fprintf(stderr, "order needs width1 width2\n");
        return 2;
    }
    mr::ENV env(dev ? std::atoi(dev) : 0, cfg ? cfg : "config.txt", lg ? lg : "log.txt");
    float speedup = 0.0f;
    mr::GRAPH<float> graph(env, gid);
    if (!graph.mat) return 1;
    std::cout << "graph_id = " << graph.graph_id << ", type = " << type;
    env.FSTR << graph.graph_id;
    try {
        if (type == "origin") {
            env.FSTR << "&origin&&&";
            graph.locality_predict(env, speedup);
        } else if (type == "degree") {
            graph.relabel_degree();
            graph.locality_predict(env, speedup);
        } else if (type == "dbg") {
            graph.dbg(env);
            graph.locality_predict(env, speedup);
        } else if (type == "rcm") {
            graph.relabel_rcm();
            graph.locality_predict(env, speedup);
        } else if (type == "rabbitorder") {
            graph.rabbitorder();
            graph.locality_predict(env, speedup);
        } else if (type == "gorder1") {
            graph.gorder1(env, 5);
            graph.locality_predict(env, speedup);
        } else if (type == "gorder5") {
            graph.gorder5(env, graph.get_width(), 1, 5);
            graph.locality_predict(env, speedup);
        } else if (type == "gorder6") {
            graph.locality_predict(env, speedup);   // on the order before pGOrder (main.cu:122-126)
            graph.gorder6(env, graph.get_width());
        } else if (type == "order") {
            graph.order(env, std::stoi(argv[3]), std::stoi(argv[4]));
        } else {
            std::fprintf(stderr, "unknown type %s\n", type.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    graph.spmv(env, "merbit");
    return 0;
}
