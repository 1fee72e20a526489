// CPU harness over the library's launch plan (gpuothello2_amd/csrc/plan.hpp):
//   plan_check N MAX_DEPTH CUS PER_CU [field=value ...] [capacity=BYTES]
// fills a go2_config with go2_config_default's values, applies the given
// integer fields, and prints the plan as "key value" lines; with capacity=,
// the plan shrunk to fit that many bytes of scratch (plan_grid_within).  Built and run
// by tests/test_launch_plan.py with g++ (no HIP needed).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "plan.hpp"

static void defaults(go2_config* c) {
  // Mirrors go2_config_default (csrc/api.cpp), which needs the HIP library.
  std::memset(c, 0, sizeof(*c));
  c->ordering = GO2_ORDER_MOBILITY;
  c->mobility_min_empties = 7;
  c->learned_min_empties = 11;
  c->block_size = 256;
  c->time_limit_ms = 600000;
  c->aspiration_width = 8;
  c->pvs_min_empties = 8;
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s N MAX_DEPTH CUS PER_CU [field=value ...]\n", argv[0]);
    return 1;
  }
  go2_config c;
  defaults(&c);
  struct Field {
    const char* name;
    int32_t* slot;
  } fields[] = {{"ordering", &c.ordering},
                {"block_size", &c.block_size},
                {"blocks_per_cu", &c.blocks_per_cu},
                {"positions_per_wave", &c.positions_per_wave},
                {"split_min_empties", &c.split_min_empties},
                {"lds_levels", &c.lds_levels},
                {"export_min_empties", &c.export_min_empties},
                {"export_low_empties", &c.export_low_empties},
                {"think_depth", &c.think_depth}};
  long long capacity = -1;
  for (int i = 5; i < argc; ++i) {
    const char* eq = std::strchr(argv[i], '=');
    if (!eq) return 2;
    if (!std::strncmp(argv[i], "capacity=", 9)) {
      capacity = std::atoll(eq + 1);
      continue;
    }
    bool known = false;
    for (auto& f : fields)
      if (std::strlen(f.name) == (size_t)(eq - argv[i]) && !std::strncmp(f.name, argv[i], eq - argv[i])) {
        *f.slot = std::atoi(eq + 1);
        known = true;
      }
    if (!known) return 2;
  }
  const go2::StackPlan s = go2::plan_stack(std::atoi(argv[2]), c);
  go2::GridPlan g;
  if (capacity >= 0) {
    const bool fits = go2::plan_grid_within(s, (size_t)std::atoll(argv[1]), c, std::atoi(argv[3]),
                                            std::atoi(argv[4]), (size_t)capacity, &g);
    std::printf("fits %d\n", (int)fits);
    if (!fits) return 0;
  } else {
    g = go2::plan_grid(s, (size_t)std::atoll(argv[1]), c, std::atoi(argv[3]), std::atoi(argv[4]));
  }
  std::printf("scratch_bytes %zu\nframe_bytes %zu\nhelp_bytes %zu\n", go2::scratch_bytes(s, g),
              go2::frame_bytes(g), go2::help_bytes(s, g));
  std::printf("think %d\ndepth %d\nmaxe %d\nlevels %d\nlds_levels %d\nlds_bytes %zu\n", (int)s.think, s.depth, s.maxe,
              s.levels, s.lds_levels, s.lds_bytes);
  std::printf("grid %zu\nwaves %zu\ntotal_lanes %zu\ng_levels %zu\nK %d\nsplit_min %d\next_min %d\next_lo %d\nhelp %d\n",
              g.grid, g.waves, g.total_lanes, g.g_levels, g.K, g.split_min, g.ext_min, g.ext_lo, (int)g.help);
  return 0;
}
