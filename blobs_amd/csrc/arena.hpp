// arena.hpp -- generational arena with the iteration and slot-reuse order of
// thunderdome 0.6 (Cargo.lock:1694-1697), which the reference uses for every
// body, collider, joint and spring set (physics.rs:11-16).
//
//  * iteration is in slot order (so "arena order" == slot order);
//  * remove() pushes the slot on a LIFO free list, insert() pops it;
//  * handles are Index::to_bits() = (generation << 32) | slot, generation 1 on
//    a fresh slot and +1 on every reuse.
#pragma once

#include <stdint.h>

#include <vector>

namespace bg {

template <class T>
class Arena {
public:
    struct Slot {
        T value;
        uint32_t gen = 0;
        bool occupied = false;
        int32_t next_free = -1;
    };

    static uint64_t make_handle(uint32_t slot, uint32_t gen) { return ((uint64_t)gen << 32) | slot; }
    static uint32_t slot_of(uint64_t h) { return (uint32_t)(h & 0xffffffffu); }
    static uint32_t gen_of(uint64_t h) { return (uint32_t)(h >> 32); }

    uint64_t insert(const T& v) {
        ++len_;
        if (first_free_ >= 0) {
            uint32_t s = (uint32_t)first_free_;
            Slot& e = slots_[s];
            first_free_ = e.next_free;
            e.gen = e.gen + 1;
            e.occupied = true;
            e.next_free = -1;
            e.value = v;
            return make_handle(s, e.gen);
        }
        Slot e;
        e.value = v;
        e.gen = 1;
        e.occupied = true;
        slots_.push_back(e);
        return make_handle((uint32_t)(slots_.size() - 1), 1);
    }

    T* get(uint64_t h) {
        uint32_t s = slot_of(h);
        if (s >= slots_.size()) return nullptr;
        Slot& e = slots_[s];
        if (!e.occupied || e.gen != gen_of(h)) return nullptr;
        return &e.value;
    }
    const T* get(uint64_t h) const { return const_cast<Arena*>(this)->get(h); }

    bool remove(uint64_t h) {
        T* p = get(h);
        if (!p) return false;
        uint32_t s = slot_of(h);
        Slot& e = slots_[s];
        e.occupied = false;
        e.value = T();
        e.next_free = first_free_;
        first_free_ = (int32_t)s;
        --len_;
        return true;
    }

    void clear() {
        slots_.clear();
        first_free_ = -1;
        len_ = 0;
    }

    uint32_t len() const { return len_; }
    uint32_t capacity_slots() const { return (uint32_t)slots_.size(); }
    bool occupied(uint32_t s) const { return s < slots_.size() && slots_[s].occupied; }
    T& at_slot(uint32_t s) { return slots_[s].value; }
    const T& at_slot(uint32_t s) const { return slots_[s].value; }
    uint64_t handle_at(uint32_t s) const { return make_handle(s, slots_[s].gen); }
    // geometric growth: repeated small bulk inserts stay linear overall
    void reserve(size_t n) {
        if (n > slots_.capacity()) slots_.reserve(n > 2 * slots_.capacity() ? n : 2 * slots_.capacity());
    }

private:
    std::vector<Slot> slots_;
    int32_t first_free_ = -1;
    uint32_t len_ = 0;
};

}  // namespace bg
