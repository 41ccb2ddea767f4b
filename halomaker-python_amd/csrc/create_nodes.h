// create_nodes.h -- host structure-tree stage (see create_nodes.cpp).
#pragma once
#include <cstdint>
#include <vector>

namespace ahop {

struct TreeParams {
    double rho_threshold = 80.0, fudge = 4.0, alphap = 1.0, epsilon = 0.0;
    int32_t nmembthresh = 20;
    double xlong = 1.0, ylong = 1.0, zlong = 1.0;
    double xlongs2 = 0.5, ylongs2 = 0.5, zlongs2 = 0.5;
    int32_t nthreads = 0;   // OpenMP threads of the stage (0: the OpenMP default)
    // the reference's moment-prefix node creation (optimize_nodes, f90:954-958): a component
    // of 1000 members or more, not within 1e-8 of a selection cut, takes its node properties
    // from per-group prefix sums in descending density instead of the member lists
    bool optimize_nodes = false;
};

// Inputs in original particle order (host memory).
struct TreeInputs {
    int64_t npart = 0;
    const double *pos = nullptr;      // (npart,3) row-major == reference pos(3,npart)
    const double *mass = nullptr;
    const double *density = nullptr;
    const int32_t *gid = nullptr;     // group ids from the local-maxima stage (0 = inactive)
    int32_t ngroups = 0;
    const double *densityg = nullptr; // [ngroups]
    const int32_t *sad_offset = nullptr;  // [ngroups+1], 0-based
    const int32_t *sad_isad = nullptr;    // 1-based neighbour group ids, decreasing saddle order
    const double *sad_rho = nullptr;
};

struct TreeOutputs {
    int32_t nnodes = 0;
    std::vector<int32_t> liste_parts;     // node id of each particle (0 = none); may come in
                                          // holding a previous buffer, which is reused
    std::vector<double> real_table;       // (8, nnodes) column-major
    std::vector<int32_t> integer_table;   // (6, nnodes) column-major
    double ms_link = 0, ms_colour = 0, ms_nodes = 0, ms_reorder = 0, ms_output = 0;   // stages
    // parallel branches taken: levels whose colour components were treated in a taskloop,
    // the components treated there, and child subtrees deferred as tasks
    int64_t taskloop_levels = 0, taskloop_components = 0, subtree_tasks = 0;
};

void create_group_tree(const TreeInputs &in, const TreeParams &p, TreeOutputs &out);

}  // namespace ahop
