// Slice frames and the mmap snapshot file (reference interface:
// src/cpp/rtm/mmap_writer.h:24-121 and the frame format of
// src/cpp/rtm/snapshot_buffer.h:18-37, snapshot_buffer.cpp:98-227).
//
// Frame (little-endian): a 128-byte header — magic u64, timestep i64, axis
// i32, index i64, dim0 i64, dim1 i64, total_size i64, forward/adjoint/ic
// offsets i64 x3, forward/adjoint/ic byte counts i64 x3, capture_timestamp
// f64, sequence_id u64, checksum u32, reserved u32 (124 packed bytes) and 4
// zero bytes — then the forward, adjoint and image-condition planes (FP64).
// The checksum is the xor of the payload's 32-bit words, rotated left by 13
// after each; a trailing partial word is xor-ed in without rotation.
//
// File (version 2 of the reference's layout): the 120-byte MmapFileHeader,
// zero-padded to header_size = 128 bytes, then length-prefixed frames
// ([i64 size][frame]) from byte 128 up to write_offset.  The reference's C++
// writer starts frames at 120 while its Python reader starts at 128; version
// 2 states the start in header_size and uses 128.  A frame that does not fit
// behind write_offset restarts the ring at 128 (ring_start counts restarts)
// and frame_count is the number of frames between 128 and write_offset.
#pragma once

#include <cstddef>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "rtm/rtm_engine.h"

namespace seismic {

constexpr uint64_t SLICE_FRAME_MAGIC = 0x534C494345303031ULL;
constexpr int64_t SLICE_FRAME_HEADER_SIZE = 128;
constexpr int64_t DEFAULT_MMAP_MAX_BYTES = 4LL * 1024LL * 1024LL * 1024LL;
constexpr int64_t MMAP_FILE_MAGIC = 0x52544D534E4150LL;
constexpr int64_t MMAP_HEADER_BYTES = 128;

// every field is 8 bytes at an 8-byte offset, so the natural layout is the
// packed one and the atomics on frame_count / write_offset / ring_start are
// aligned (the mapping starts on a page)
struct MmapFileHeader {
    int64_t magic;
    int64_t header_size;
    int64_t file_size_bytes;
    int64_t frame_count;
    int64_t write_offset;
    int64_t ring_start;
    int64_t ring_capacity;
    int64_t reserved[8];
};
static_assert(sizeof(MmapFileHeader) == 120, "MmapFileHeader layout");
static_assert(offsetof(MmapFileHeader, frame_count) == 24 && offsetof(MmapFileHeader, ring_start) == 40,
              "MmapFileHeader layout");

uint32_t slice_checksum(const uint8_t* data, size_t n);
// serialize_slice / deserialize_slice of the reference's snapshot buffer;
// deserialize returns false unless magic, sizes and checksum agree
void serialize_slice(const RTMEngine::CapturedSlice& s, uint64_t sequence_id, std::vector<uint8_t>& out);
bool deserialize_slice(const uint8_t* data, size_t n, RTMEngine::CapturedSlice& out, uint64_t* sequence_id);

class MmapSnapshotWriter {
public:
    MmapSnapshotWriter() = default;
    ~MmapSnapshotWriter();
    MmapSnapshotWriter(const MmapSnapshotWriter&) = delete;
    MmapSnapshotWriter& operator=(const MmapSnapshotWriter&) = delete;

    // creates (truncates) the file at max_bytes and maps it
    bool open(const std::string& filename, int64_t max_bytes = DEFAULT_MMAP_MAX_BYTES);
    void close();
    bool is_open() const { return base_ != nullptr; }

    // false when closed, empty, or larger than the frame region
    bool write_frame(const uint8_t* frame_data, int64_t frame_size);
    bool write_slice(const RTMEngine::CapturedSlice& s);

    int64_t frame_count() const { return header()->frame_count; }
    int64_t write_offset() const { return header()->write_offset; }
    int64_t used_bytes() const { return write_offset(); }
    int64_t capacity_bytes() const { return max_bytes_; }

    struct Stats {
        int64_t frames_written;
        int64_t frames_dropped_buffer_full;
        int64_t total_bytes_written;
        int64_t mmap_cycles;
    };
    Stats stats() const;
    std::string filename() const { return filename_; }

    static bool read_frame(const std::string& filename, int64_t frame_index, std::vector<uint8_t>& out_data);
    static int64_t count_frames(const std::string& filename);

    MmapFileHeader* header() { return reinterpret_cast<MmapFileHeader*>(base_); }
    const MmapFileHeader* header() const { return reinterpret_cast<const MmapFileHeader*>(base_); }

private:
    std::string filename_;
    int64_t max_bytes_ = 0;
    void* base_ = nullptr;
    int fd_ = -1;
    uint64_t next_sequence_ = 0;
    std::atomic<int64_t> frames_written_{0}, frames_dropped_{0}, bytes_written_{0}, cycles_{0};
};

}  // namespace seismic
