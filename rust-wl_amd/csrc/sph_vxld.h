// VXLD v3 frame dumps: the binary format the reference's demos write and its
// renderer reads (crates/vxl-phys/examples/dam_break.rs:14-44, header at
// :84-110, frames at :120-128; body records as showcase.rs:311-360; reader
// scripts/render_demo.py:39-100). All values little-endian.
//
//   header  "VXLD" u32 version=3, u32 ticks, u32 ticks_per_frame, f32 dt,
//           u32 voxel dims[3], f32 voxel origin[3], f32 voxel step,
//           u32 splats (0 here), u32 meshes (0 here), u32 fluid systems
//   frame   u32 tick, f32 ms, u32 bodies,
//           per body: u8 kind (0 = box), u8 awake, f32 pos[3], f32 rot[4], f32 half[3]
//           voxel occupancy bits (none here: dims are 0),
//           per fluid system: u32 count, f32 xyz per particle
#ifndef VXL_SPH_VXLD_H
#define VXL_SPH_VXLD_H
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace vxl {

struct VxldBody {
    uint8_t awake;   // 1 for dynamic bodies
    float pos[3];
    float rot[4];    // x, y, z, w
    float half[3];
};

class VxldWriter {
public:
    VxldWriter(const std::string &path, uint32_t ticks, uint32_t ticks_per_frame, float dt, uint32_t nfluid);
    ~VxldWriter();
    VxldWriter(const VxldWriter &) = delete;
    VxldWriter &operator=(const VxldWriter &) = delete;
    // fluids[k] = (xyz of count particles, count); exactly nfluid systems.
    void frame(uint32_t tick, float ms, const std::vector<VxldBody> &bodies,
               const std::vector<std::pair<const float *, uint32_t>> &fluids);
    void close();
    uint32_t nfluid() const { return nfluid_; }

private:
    void put(const void *p, size_t n);
    template <class T>
    void put(T v) {
        put(&v, sizeof(T));
    }
    std::FILE *f_ = nullptr;
    uint32_t nfluid_ = 0;
};

}  // namespace vxl
#endif
