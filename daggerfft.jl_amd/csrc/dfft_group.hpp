// dfft_group.hpp -- grouped launches of independent axis passes (see dfft_group.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "dfft_plan.hpp"
#include "dfft_types.hpp"

namespace dfft {

constexpr int GROUP_MAX_SLOTS = 256;

// One axis pass over a slice of two pointer slots, cut into `ntiles` tiles.
struct GroupItem {
    AxisPass ap;              // geometry + schedule; ap.in/ap.out unused
    int prec;
    int in_slot, out_slot;
    int64_t in_off, out_off;  // byte offsets into the slots
    int64_t ntiles;
    int64_t first_block;      // first block of this item in its launch
    int threads;              // threads a tile needs (<= the launch's workgroup size)
    size_t lds;               // tile LDS bytes
    int gfix;                 // >= 0: compile-time schedule gfix_entry(gfix); -1: runtime engine
};

// Compile-time C2C schedules the grouped kernel runs in its 256-thread
// launches (see dfft_group.hip), one entry per (length, precision, layout).
struct GFixEntry {
    int L, prec, contig, emax, split, tpl, S;
    const int *R;
};
constexpr int GFIX_THREADS = 256;
// entry of (prec, contig, L), or -1
int gfix_lookup(int prec, int contig, int L);
const GFixEntry &gfix_entry(int id);

struct GroupExec;

// levels[l] = the items of launch l (independent of each other)
int group_create(GroupExec **out, const std::vector<std::vector<GroupItem>> &levels);
int group_launch(const GroupExec *g, void *const *slots, int nslots, hipStream_t st);
void group_destroy(GroupExec *g);
struct GroupGraph;
GroupGraph *group_graph_new();
void group_graph_destroy(GroupGraph *gg);
// the group's launches as one hipGraph replay (recorded on first use and
// whenever the slot pointers change)
int group_launch_graph(const GroupExec *g, GroupGraph *gg, void *const *slots, int nslots, hipStream_t st);
// the whole group as one cooperative launch with a grid barrier between
// levels: prepare once (every level must run 256-thread tiles; sizes the
// grid to the workgroups the device holds at once), then launch
int group_persistent_prepare(GroupExec *g);
int group_persistent_grid(const GroupExec *g);
int group_launch_persistent(const GroupExec *g, void *const *slots, int nslots, hipStream_t st);
int group_num_launches(const GroupExec *g);
int64_t group_blocks(const GroupExec *g);
GroupItem make_group_item(const PassPlan &pp, int in_slot, int64_t in_off, int out_slot, int64_t out_off);

}  // namespace dfft
