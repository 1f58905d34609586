// Host-side volume file readers (RAW slice folders and TIFF stacks) shared by
// the device ingest (s2hip_volin.hip) and the host-only read entry points.
#pragma once
#include <cstdint>
#include <functional>
#include <string>

#include "../../include/s2hip.h"

namespace s2h {
namespace volio {

// Receives the slices of a volume in z order.  slot(bytes) returns where the
// next slice (width * height samples, bytes long) is to be written; commit()
// is called once it is complete.  Either returns false to stop with an error
// the callee has already recorded in err.
struct SliceSink {
    std::function<uint8_t *(uint64_t bytes, std::string &err)> slot;
    std::function<bool(std::string &err)> commit;
};

// Dimensions and sample layout of the volume src describes, after the slice
// range is applied.  The staged byte order is the RAW spec's, or the first TIFF
// file's (later TIFF pages are converted to it).
bool probe(const s2h_volume_source &src, s2h_volume_info &info, std::string &err);

// Decodes every slice of the range into sink, in the layout probe reported.
bool stream(const s2h_volume_source &src, const s2h_volume_info &info, const SliceSink &sink, std::string &err);

}  // namespace volio
}  // namespace s2h
