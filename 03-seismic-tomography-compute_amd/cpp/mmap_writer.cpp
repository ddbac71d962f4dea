// Slice frames and the mmap snapshot file (reference: src/cpp/rtm/
// mmap_writer.cpp:34-331, snapshot_buffer.cpp:98-227).  Host-side file I/O of
// slices captured by RTMEngine::extract_slices; the layout is documented in
// rtm/mmap_writer.h.
#include "rtm/mmap_writer.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>

namespace seismic {

namespace {

#pragma pack(push, 1)
struct FrameHeader {
    uint64_t magic;
    int64_t timestep;
    int32_t axis;
    int64_t index, dim0, dim1, total_size;
    int64_t fwd_off, adj_off, ic_off, fwd_bytes, adj_bytes, ic_bytes;
    double capture_timestamp;
    uint64_t sequence_id;
    uint32_t checksum, reserved;
};
#pragma pack(pop)
static_assert(sizeof(FrameHeader) == 124, "slice frame header layout");

template <typename T>
T load(const uint8_t* p) {
    T v;
    std::memcpy(&v, p, sizeof v);
    return v;
}

// read-only view of a whole file
struct FileView {
    const uint8_t* base = nullptr;
    size_t size = 0;
    int fd = -1;
    explicit FileView(const std::string& name) {
        fd = ::open(name.c_str(), O_RDONLY);
        if (fd < 0) return;
        struct stat st;
        if (::fstat(fd, &st) != 0 || st.st_size < MMAP_HEADER_BYTES) return;
        void* p = ::mmap(nullptr, static_cast<size_t>(st.st_size), PROT_READ, MAP_PRIVATE, fd, 0);
        if (p == MAP_FAILED) return;
        base = static_cast<const uint8_t*>(p);
        size = static_cast<size_t>(st.st_size);
    }
    ~FileView() {
        if (base) ::munmap(const_cast<uint8_t*>(base), size);
        if (fd >= 0) ::close(fd);
    }
    bool valid() const {
        return base && load<int64_t>(base) == MMAP_FILE_MAGIC && load<int64_t>(base + 8) >= 120 &&
               load<int64_t>(base + 8) <= static_cast<int64_t>(size);
    }
};

}  // namespace

uint32_t slice_checksum(const uint8_t* data, size_t n) {
    uint32_t s = 0;
    const size_t words = n / 4;
    for (size_t i = 0; i < words; ++i) {
        s ^= load<uint32_t>(data + 4 * i);
        s = (s << 13) | (s >> 19);
    }
    uint32_t tail = 0;
    for (size_t i = 4 * words; i < n; ++i) tail |= static_cast<uint32_t>(data[i]) << (8 * (i - 4 * words));
    return s ^ tail;
}

void serialize_slice(const RTMEngine::CapturedSlice& s, uint64_t sequence_id, std::vector<uint8_t>& out) {
    const int64_t fb = static_cast<int64_t>(s.forward_slice.size() * sizeof(double));
    const int64_t ab = static_cast<int64_t>(s.adjoint_slice.size() * sizeof(double));
    const int64_t ib = static_cast<int64_t>(s.ic_slice.size() * sizeof(double));
    out.assign(static_cast<size_t>(SLICE_FRAME_HEADER_SIZE + fb + ab + ib), 0);
    uint8_t* p = out.data() + SLICE_FRAME_HEADER_SIZE;
    if (fb) std::memcpy(p, s.forward_slice.data(), static_cast<size_t>(fb));
    if (ab) std::memcpy(p + fb, s.adjoint_slice.data(), static_cast<size_t>(ab));
    if (ib) std::memcpy(p + fb + ab, s.ic_slice.data(), static_cast<size_t>(ib));
    FrameHeader h{};
    h.magic = SLICE_FRAME_MAGIC;
    h.timestep = s.timestep;
    h.axis = static_cast<int32_t>(s.axis);
    h.index = s.index;
    h.dim0 = s.dim0;
    h.dim1 = s.dim1;
    h.total_size = static_cast<int64_t>(out.size());
    h.fwd_off = SLICE_FRAME_HEADER_SIZE;
    h.adj_off = SLICE_FRAME_HEADER_SIZE + fb;
    h.ic_off = SLICE_FRAME_HEADER_SIZE + fb + ab;
    h.fwd_bytes = fb;
    h.adj_bytes = ab;
    h.ic_bytes = ib;
    h.capture_timestamp = s.capture_timestamp;
    h.sequence_id = sequence_id;
    h.checksum = slice_checksum(p, static_cast<size_t>(fb + ab + ib));
    std::memcpy(out.data(), &h, sizeof h);   // bytes 124-127 stay zero
}

bool deserialize_slice(const uint8_t* data, size_t n, RTMEngine::CapturedSlice& out, uint64_t* sequence_id) {
    if (!data || n < static_cast<size_t>(SLICE_FRAME_HEADER_SIZE)) return false;
    FrameHeader h;
    std::memcpy(&h, data, sizeof h);
    if (h.magic != SLICE_FRAME_MAGIC || h.total_size < SLICE_FRAME_HEADER_SIZE ||
        h.total_size > static_cast<int64_t>(n))
        return false;
    const int64_t ts = h.total_size;
    if (slice_checksum(data + SLICE_FRAME_HEADER_SIZE, static_cast<size_t>(ts - SLICE_FRAME_HEADER_SIZE)) !=
        h.checksum)
        return false;
    const int64_t offs[3] = {h.fwd_off, h.adj_off, h.ic_off}, lens[3] = {h.fwd_bytes, h.adj_bytes, h.ic_bytes};
    for (int k = 0; k < 3; ++k)
        if (offs[k] < 0 || lens[k] < 0 || lens[k] % 8 || offs[k] > ts || lens[k] > ts - offs[k]) return false;
    out.timestep = h.timestep;
    out.axis = static_cast<SliceAxis>(h.axis);
    out.index = h.index;
    out.dim0 = h.dim0;
    out.dim1 = h.dim1;
    out.capture_timestamp = h.capture_timestamp;
    std::vector<double>* dst[3] = {&out.forward_slice, &out.adjoint_slice, &out.ic_slice};
    for (int k = 0; k < 3; ++k) {
        dst[k]->resize(static_cast<size_t>(lens[k] / 8));
        if (lens[k]) std::memcpy(dst[k]->data(), data + offs[k], static_cast<size_t>(lens[k]));
    }
    if (sequence_id) *sequence_id = h.sequence_id;
    return true;
}

MmapSnapshotWriter::~MmapSnapshotWriter() { close(); }

bool MmapSnapshotWriter::open(const std::string& filename, int64_t max_bytes) {
    close();
    if (max_bytes < MMAP_HEADER_BYTES + 8) return false;
    fd_ = ::open(filename.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (fd_ < 0) return false;
    if (::ftruncate(fd_, max_bytes) != 0) {
        ::close(fd_);
        fd_ = -1;
        return false;
    }
    void* p = ::mmap(nullptr, static_cast<size_t>(max_bytes), PROT_READ | PROT_WRITE, MAP_SHARED, fd_, 0);
    if (p == MAP_FAILED) {
        ::close(fd_);
        fd_ = -1;
        return false;
    }
    base_ = p;
    filename_ = filename;
    max_bytes_ = max_bytes;
    std::memset(base_, 0, static_cast<size_t>(MMAP_HEADER_BYTES));
    MmapFileHeader* h = header();
    h->magic = MMAP_FILE_MAGIC;
    h->header_size = MMAP_HEADER_BYTES;
    h->file_size_bytes = max_bytes;
    h->write_offset = MMAP_HEADER_BYTES;
    cycles_.fetch_add(1, std::memory_order_relaxed);
    return true;
}

void MmapSnapshotWriter::close() {
    if (!base_) return;
    ::msync(base_, static_cast<size_t>(max_bytes_), MS_SYNC);
    ::munmap(base_, static_cast<size_t>(max_bytes_));
    ::close(fd_);
    base_ = nullptr;
    fd_ = -1;
}

bool MmapSnapshotWriter::write_frame(const uint8_t* frame_data, int64_t frame_size) {
    if (!base_ || !frame_data || frame_size <= 0 || frame_size > max_bytes_ - MMAP_HEADER_BYTES - 8) {
        frames_dropped_.fetch_add(1, std::memory_order_relaxed);
        return false;
    }
    MmapFileHeader* h = header();
    int64_t off = __atomic_load_n(&h->write_offset, __ATOMIC_ACQUIRE);
    if (off + 8 + frame_size > max_bytes_) {
        // restart the ring: readers see no frames until the new one is published
        __atomic_store_n(&h->frame_count, 0, __ATOMIC_RELEASE);
        __atomic_store_n(&h->write_offset, MMAP_HEADER_BYTES, __ATOMIC_RELEASE);
        __atomic_add_fetch(&h->ring_start, 1, __ATOMIC_ACQ_REL);
        cycles_.fetch_add(1, std::memory_order_relaxed);
        off = MMAP_HEADER_BYTES;
    }
    uint8_t* dst = static_cast<uint8_t*>(base_) + off;
    std::memcpy(dst, &frame_size, sizeof frame_size);
    std::memcpy(dst + 8, frame_data, static_cast<size_t>(frame_size));
    __atomic_store_n(&h->write_offset, off + 8 + frame_size, __ATOMIC_RELEASE);
    __atomic_add_fetch(&h->frame_count, 1, __ATOMIC_ACQ_REL);
    frames_written_.fetch_add(1, std::memory_order_relaxed);
    bytes_written_.fetch_add(8 + frame_size, std::memory_order_relaxed);
    return true;
}

bool MmapSnapshotWriter::write_slice(const RTMEngine::CapturedSlice& s) {
    std::vector<uint8_t> buf;
    serialize_slice(s, next_sequence_++, buf);
    return write_frame(buf.data(), static_cast<int64_t>(buf.size()));
}

MmapSnapshotWriter::Stats MmapSnapshotWriter::stats() const {
    return Stats{frames_written_.load(), frames_dropped_.load(), bytes_written_.load(), cycles_.load()};
}

bool MmapSnapshotWriter::read_frame(const std::string& filename, int64_t frame_index, std::vector<uint8_t>& out) {
    FileView v(filename);
    if (!v.valid() || frame_index < 0) return false;
    int64_t off = load<int64_t>(v.base + 8);
    const int64_t end = std::min<int64_t>(load<int64_t>(v.base + 32), static_cast<int64_t>(v.size));
    for (int64_t k = 0; off + 8 <= end; ++k) {
        const int64_t n = load<int64_t>(v.base + off);
        off += 8;
        if (n <= 0 || n > end - off) return false;
        if (k == frame_index) {
            out.assign(v.base + off, v.base + off + n);
            return true;
        }
        off += n;
    }
    return false;
}

int64_t MmapSnapshotWriter::count_frames(const std::string& filename) {
    FileView v(filename);
    return v.valid() ? load<int64_t>(v.base + 24) : -1;
}

}  // namespace seismic
