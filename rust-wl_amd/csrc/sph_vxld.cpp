// VXLD v3 writer (see sph_vxld.h). The host is little-endian (x86-64), so
// values are written in their in-memory byte order.
#include "sph_vxld.h"

#include <stdexcept>

namespace vxl {

VxldWriter::VxldWriter(const std::string &path, uint32_t ticks, uint32_t ticks_per_frame, float dt, uint32_t nfluid)
    : nfluid_(nfluid) {
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) throw std::runtime_error("cannot open " + path + " for writing");
    put("VXLD", 4);
    put<uint32_t>(3u);
    put<uint32_t>(ticks);
    put<uint32_t>(ticks_per_frame);
    put<float>(dt);
    for (int a = 0; a < 3; ++a) put<uint32_t>(0u);  // no voxel volume
    for (int a = 0; a < 4; ++a) put<float>(0.0f);   // its origin and step
    put<uint32_t>(0u);                              // splats
    put<uint32_t>(0u);                              // meshes
    put<uint32_t>(nfluid);
}

VxldWriter::~VxldWriter() {
    if (f_) std::fclose(f_);
}

void VxldWriter::put(const void *p, size_t n) {
    if (std::fwrite(p, 1, n, f_) != n) throw std::runtime_error("VXLD write failed");
}

void VxldWriter::frame(uint32_t tick, float ms, const std::vector<VxldBody> &bodies,
                       const std::vector<std::pair<const float *, uint32_t>> &fluids) {
    if (!f_) throw std::logic_error("VXLD writer is closed");
    if (fluids.size() != nfluid_) throw std::invalid_argument("frame fluid systems differ from the header");
    put<uint32_t>(tick);
    put<float>(ms);
    put<uint32_t>((uint32_t)bodies.size());
    for (const VxldBody &b : bodies) {
        put<uint8_t>(0u);  // box
        put<uint8_t>(b.awake);
        put(b.pos, 12);
        put(b.rot, 16);
        put(b.half, 12);
    }
    // voxel occupancy: none (0 cells)
    for (const auto &fl : fluids) {
        put<uint32_t>(fl.second);
        if (fl.second) put(fl.first, 12ull * fl.second);
    }
}

void VxldWriter::close() {
    if (f_ && std::fclose(f_) != 0) {
        f_ = nullptr;
        throw std::runtime_error("VXLD close failed");
    }
    f_ = nullptr;
}

}  // namespace vxl
