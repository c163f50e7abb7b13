// Build schedule of the persistent search: per variant, size classes that give the waves per CU and the seeds per
// lane of the leaf and split windows for batches up to a key count. The autotuner (shgpu_autotune_class) measures a
// class on a sample; the table persists as JSON, which the engine loads when it is created. Host code only.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/shockhash_amd.h"

namespace shgpu {

struct Schedule {
    std::vector<shgpu_schedule_class> classes;  // sorted by (variant, max_keys); every variant has at least one
    std::string source = "built-in";            // the file it came from

    // The hand-measured table (ShockHash2 leaf 64 and ShockHash leaf 30 kernel times at 4 to 24 waves per CU),
    // used when no schedule file is found.
    static Schedule builtin();
    // The class of the variant's first max_keys >= n; the variant's largest class takes larger batches.
    const shgpu_schedule_class &lookup(uint32_t variant, uint64_t n) const;
    // Replaces the class of the same variant and max_keys, or adds it.
    void put(const shgpu_schedule_class &c);
    // Replaces the classes of every variant that `cls` names. SHGPU_ERR_ARG on an invalid class.
    int replace(const std::vector<shgpu_schedule_class> &cls, std::string &err);
    std::string to_json(const std::string &arch) const;
};

// Checks one class: known variant, waves 1..32, seeds per lane 1, 2 or 4, max_keys > 0.
int check_class(const shgpu_schedule_class &c, std::string &err);
// Parses the JSON form written by Schedule::to_json ({"arch": ..., "classes": [{"variant": ..., ...}, ...]}):
// every object in "classes" gives one class; unknown keys are ignored, missing ones are an error.
int parse_schedule_json(const std::string &text, std::vector<shgpu_schedule_class> &out, std::string &err);

}  // namespace shgpu
