// Heap for buffers that come and go at every step of a server: chunks taken
// from a backend (hipMalloc on the device) are carved into blocks best-fit,
// a freed block merges with its free neighbours in the same chunk, and a
// chunk that is entirely free goes back to the backend only when the chunks
// together exceed a set size.  Steady use therefore makes no backend call.
//
// Host code only; not thread-safe (one owner).  The backend is a type with
//     void *get(uint64_t bytes);   // nullptr when it has no memory
//     void put(void *p);
// so that the same heap runs over malloc in the CPU tests.
#pragma once
#include <stdint.h>
#include <algorithm>
#include <iterator>
#include <map>
#include <vector>

namespace rpx {

template <class Backend>
class BlockHeap {
  public:
    static constexpr uint64_t kAlign = 256;            // block sizes and offsets
    static constexpr uint64_t kChunkAlign = 2ull << 20;

    explicit BlockHeap(Backend be = Backend()) : be_(be) {}
    BlockHeap(const BlockHeap &) = delete;
    BlockHeap &operator=(const BlockHeap &) = delete;
    ~BlockHeap() { release_all(); }

    // keep: chunk bytes above which entirely free chunks are released;
    // min_chunk: smallest chunk requested from the backend
    void set_limits(uint64_t keep, uint64_t min_chunk) {
        keep_ = keep;
        min_chunk_ = min_chunk;
    }

    // A block of at least `bytes` bytes, kAlign-aligned relative to its
    // chunk; nullptr when the backend cannot provide a chunk for it.
    void *alloc(uint64_t bytes) {
        const uint64_t need = round_up(std::max<uint64_t>(bytes, 1), kAlign);
        auto it = free_by_size_.lower_bound(need);
        if (it == free_by_size_.end()) {
            if (!grow(need)) return nullptr;
            it = free_by_size_.lower_bound(need);
            if (it == free_by_size_.end()) return nullptr;
        }
        char *p = it->second;
        auto fa = free_by_addr_.find(p);
        Blk b = fa->second;
        free_by_size_.erase(it);
        free_by_addr_.erase(fa);
        if (b.size - need >= kAlign) {
            insert_free(p + need, b.size - need, b.chunk);
            b.size = need;
        }
        used_[p] = b;
        used_bytes_ += b.size;
        return p;
    }

    // Give a block back; false for a pointer the heap did not hand out.
    bool free(void *ptr) {
        char *p = (char *)ptr;
        auto u = used_.find(p);
        if (u == used_.end()) return false;
        const Blk b = u->second;
        used_.erase(u);
        used_bytes_ -= b.size;
        char *start = p;
        uint64_t size = b.size;
        auto nx = free_by_addr_.find(p + b.size);
        if (nx != free_by_addr_.end() && nx->second.chunk == b.chunk) {
            size += nx->second.size;
            erase_free(nx);
        }
        auto pv = free_by_addr_.lower_bound(p);
        if (pv != free_by_addr_.begin()) {
            --pv;
            if (pv->second.chunk == b.chunk && pv->first + pv->second.size == p) {
                start = pv->first;
                size += pv->second.size;
                erase_free(pv);
            }
        }
        insert_free(start, size, b.chunk);
        if (chunk_bytes_ > keep_) release_free_chunks(keep_);
        return true;
    }

    // Release entirely free chunks, largest first, until at most `keep`
    // chunk bytes remain or none is entirely free.
    void release_free_chunks(uint64_t keep) {
        if (chunk_bytes_ <= keep) return;
        std::vector<std::pair<uint64_t, char *>> whole;
        for (const auto &c : chunks_) {
            auto f = free_by_addr_.find(c.first);
            if (f != free_by_addr_.end() && f->second.size == c.second) whole.push_back({c.second, c.first});
        }
        std::sort(whole.begin(), whole.end(), [](const std::pair<uint64_t, char *> &a,
                                                 const std::pair<uint64_t, char *> &b) { return a.first > b.first; });
        for (const auto &w : whole) {
            if (chunk_bytes_ <= keep) break;
            erase_free(free_by_addr_.find(w.second));
            chunks_.erase(w.second);
            chunk_bytes_ -= w.first;
            be_.put(w.second);
        }
    }

    // Every chunk back to the backend (blocks still handed out included).
    void release_all() {
        for (const auto &c : chunks_) be_.put(c.first);
        chunks_.clear();
        free_by_size_.clear();
        free_by_addr_.clear();
        used_.clear();
        chunk_bytes_ = used_bytes_ = 0;
    }

    uint64_t chunk_bytes() const { return chunk_bytes_; }
    uint64_t used_bytes() const { return used_bytes_; }
    size_t n_chunks() const { return chunks_.size(); }
    size_t n_free_blocks() const { return free_by_addr_.size(); }

    // Consistency of the tables (for tests): blocks tile every chunk
    // exactly, no two free blocks of a chunk are adjacent, and the size and
    // address tables agree.
    bool check() const {
        if (free_by_size_.size() != free_by_addr_.size()) return false;
        for (const auto &f : free_by_size_) {
            auto a = free_by_addr_.find(f.second);
            if (a == free_by_addr_.end() || a->second.size != f.first) return false;
        }
        std::map<char *, std::pair<Blk, bool>> all;
        for (const auto &f : free_by_addr_) all[f.first] = {f.second, true};
        for (const auto &u : used_) {
            if (all.count(u.first)) return false;
            all[u.first] = {u.second, false};
        }
        uint64_t total = 0, used = 0;
        auto it = all.begin();
        for (const auto &c : chunks_) {
            char *at = c.first;
            bool prev_free = false;
            while (it != all.end() && it->first < c.first + c.second) {
                if (it->first != at || it->second.first.chunk != c.first) return false;
                if (it->second.second && prev_free) return false;
                prev_free = it->second.second;
                if (!it->second.second) used += it->second.first.size;
                at += it->second.first.size;
                ++it;
            }
            if (at != c.first + c.second) return false;
            total += c.second;
        }
        return it == all.end() && total == chunk_bytes_ && used == used_bytes_;
    }

  private:
    struct Blk {
        uint64_t size;
        char *chunk;
    };

    static uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

    void insert_free(char *p, uint64_t size, char *chunk) {
        free_by_addr_[p] = Blk{size, chunk};
        free_by_size_.emplace(size, p);
    }

    void erase_free(typename std::map<char *, Blk>::iterator a) {
        auto r = free_by_size_.equal_range(a->second.size);
        for (auto s = r.first; s != r.second; ++s)
            if (s->second == a->first) {
                free_by_size_.erase(s);
                break;
            }
        free_by_addr_.erase(a);
    }

    // A new chunk for a request no free block serves.  Entirely free chunks
    // (all too small for it) are released first when the new chunk would take
    // the total past keep; when the backend refuses, every entirely free
    // chunk goes back and the chunk is asked for again, then at the exact
    // size of the request.
    bool grow(uint64_t need) {
        uint64_t size = round_up(std::max(need, min_chunk_), kChunkAlign);
        if (chunk_bytes_ + size > keep_) release_free_chunks(keep_ > size ? keep_ - size : 0);
        char *p = (char *)be_.get(size);
        if (!p) {
            release_free_chunks(0);
            p = (char *)be_.get(size);
        }
        if (!p && size > need) {
            size = need;
            p = (char *)be_.get(size);
        }
        if (!p) return false;
        chunks_[p] = size;
        chunk_bytes_ += size;
        insert_free(p, size, p);
        return true;
    }

    Backend be_;
    std::map<char *, uint64_t> chunks_;            // base -> bytes
    std::multimap<uint64_t, char *> free_by_size_;
    std::map<char *, Blk> free_by_addr_;
    std::map<char *, Blk> used_;
    uint64_t chunk_bytes_ = 0, used_bytes_ = 0;
    uint64_t keep_ = ~0ull, min_chunk_ = kChunkAlign;
};

}  // namespace rpx
