// setops.hpp -- C++ host API over the MI355X intersection library (include/setops_hip.h).
//
// The reference's slice-in / visitor-out Rust API, restated for C++ callers of the GPU entries:
//   Visitor contract (setops/src/visitor.rs:15-48): visit(x) once per result, ascending;
//     Counter counts (BSR: popcount of the state, visitor.rs:399-403), VecWriter collects,
//     SliceWriter fills a caller buffer; clear() is the Clearable trait.
//   Intersect2 / Intersect2C / IntersectK and the runners run_2set, run_2set_c, run_kset,
//     run_2set_bsr (setops/src/intersect.rs:40-90).
//   svs_generic / run_svs (setops/src/intersect/svs.rs:50-73, 103-116).
//   BsrVec / BsrRef (setops/src/bsr.rs:48-164).
//   The name registry (benchmark/src/timer.rs:75-276) and Harness::time (benchmark/src/timer/
//     harness.rs:35-62: warm up for a duration, then one timed run).
// An entry runs on the GPU through the C ABI (one launch sequence per call); visitor calls replay
// the returned buffer, as run_2set_c does.  Errors (the C ABI's negative status codes) are thrown
// as setops::Error with the library's message.  Header-only; link libsetops_hip.so.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/setops_hip.h"

namespace setops {

// ---- errors -----------------------------------------------------------------------------------------

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline int check(int rc)
{
    if (rc < 0) {
        const char *m = setops_hip_last_error();
        throw Error(rc, m ? m : "");
    }
    return rc;
}

// ---- slices ---------------------------------------------------------------------------------------

struct Slice {
    const uint32_t *ptr = nullptr;
    size_t len = 0;
    Slice() = default;
    Slice(const uint32_t *p, size_t n) : ptr(p), len(n) {}
    Slice(const std::vector<uint32_t> &v) : ptr(v.data()), len(v.size()) {}  // NOLINT: implicit like &[T]
    const uint32_t *begin() const { return ptr; }
    const uint32_t *end() const { return ptr + len; }
    size_t size() const { return len; }
};

// ---- visitors (visitor.rs) --------------------------------------------------------------------------

class Counter {
public:
    void visit(uint32_t) { n_++; }
    void visit_bsr(uint32_t, uint32_t state) { n_ += (size_t)__builtin_popcount(state); }
    size_t count() const { return n_; }
    void clear() { n_ = 0; }

private:
    size_t n_ = 0;
};

class VecWriter {
public:
    VecWriter() = default;
    static VecWriter with_capacity(size_t n)
    {
        VecWriter w;
        w.v_.reserve(n);
        return w;
    }
    void visit(uint32_t x) { v_.push_back(x); }
    void clear() { v_.clear(); }
    const std::vector<uint32_t> &as_ref() const { return v_; }
    std::vector<uint32_t> into() { return std::move(v_); }

private:
    std::vector<uint32_t> v_;
};

// Writes into a caller-owned buffer; like the reference's SliceWriter it must be large enough.
class SliceWriter {
public:
    SliceWriter(uint32_t *data, size_t cap) : d_(data), cap_(cap) {}
    void visit(uint32_t x)
    {
        if (n_ >= cap_) throw Error(SETOPS_ERR_CAPACITY, "SliceWriter is full");
        d_[n_++] = x;
    }
    void clear() { n_ = 0; }
    size_t position() const { return n_; }
    Slice as_ref() const { return Slice(d_, n_); }

private:
    uint32_t *d_;
    size_t cap_, n_ = 0;
};

// ---- BSR (bsr.rs) -----------------------------------------------------------------------------------

constexpr uint32_t BSR_WIDTH = 32;
constexpr uint32_t BSR_SHIFT = 5;
constexpr uint32_t BSR_MASK = 31;

struct BsrRef {
    const uint32_t *bases = nullptr;
    const uint32_t *states = nullptr;
    size_t len = 0;
};

struct BsrVec {
    std::vector<uint32_t> bases, states;

    // one (base = x >> 5, state bit x & 31) word per distinct base of an ascending input
    static BsrVec from_sorted(Slice v)
    {
        BsrVec r;
        for (uint32_t x : v) {
            const uint32_t b = x >> BSR_SHIFT, bit = 1u << (x & BSR_MASK);
            if (!r.bases.empty() && r.bases.back() == b) r.states.back() |= bit;
            else {
                r.bases.push_back(b);
                r.states.push_back(bit);
            }
        }
        return r;
    }
    std::vector<uint32_t> to_sorted_set() const
    {
        std::vector<uint32_t> out;
        for (size_t i = 0; i < bases.size(); i++)
            for (uint32_t s = states[i]; s; s &= s - 1) out.push_back((bases[i] << BSR_SHIFT) | (uint32_t)__builtin_ctz(s));
        return out;
    }
    void visit_bsr(uint32_t base, uint32_t state)
    {
        bases.push_back(base);
        states.push_back(state);
    }
    void clear()
    {
        bases.clear();
        states.clear();
    }
    size_t len() const { return bases.size(); }
    BsrRef as_ref() const { return BsrRef{bases.data(), states.data(), bases.size()}; }
};

// ---- registry and entries (timer.rs name table) -----------------------------------------------------

class Entry {
public:
    Entry() = default;
    // Id of a registered GPU entry, e.g. "hip_galloping"; throws for an unknown name.
    static Entry lookup(const std::string &name)
    {
        Entry e;
        e.id_ = check(setops_hip_algo_lookup(name.c_str()));
        e.name_ = name;
        e.flags_ = setops_hip_algo_flags(e.id_);
        return e;
    }
    int id() const { return id_; }
    const std::string &name() const { return name_; }
    bool is_twoset() const { return flags_ & SETOPS_ALGO_TWOSET; }
    bool is_bsr() const { return flags_ & SETOPS_ALGO_BSR; }
    bool is_kset() const { return flags_ & SETOPS_ALGO_KSET; }

    // Intersect2C shape (intersect.rs:41): fills `out` (capacity >= min sizes), returns the count.
    size_t intersect_c(Slice a, Slice b, uint32_t *out, size_t out_cap) const
    {
        size_t n = 0;
        check(setops_hip_intersect2(id_, a.ptr, a.len, b.ptr, b.len, out, out_cap, &n));
        return n;
    }
    // Intersect2 shape (intersect.rs:40): visits the ascending intersection.
    template <class V>
    void operator()(Slice a, Slice b, V &visitor) const
    {
        std::vector<uint32_t> buf(std::max<size_t>(1, std::min(a.len, b.len)));
        const size_t n = intersect_c(a, b, buf.data(), buf.size());
        for (size_t i = 0; i < n; i++) visitor.visit(buf[i]);
    }
    // Intersect2Bsr shape (bsr.rs:16): visits (base, state) words, bases ascending.
    template <class V>
    void bsr(BsrRef a, BsrRef b, V &visitor) const
    {
        const size_t cap = std::max<size_t>(1, std::min(a.len, b.len));
        std::vector<uint32_t> ob(cap), os(cap);
        size_t words = 0;
        uint64_t elems = 0;
        check(setops_hip_intersect2_bsr(id_, a.bases, a.states, a.len, b.bases, b.states, b.len, ob.data(), os.data(),
                                        cap, &words, &elems));
        for (size_t i = 0; i < words; i++) visitor.visit_bsr(ob[i], os[i]);
    }
    // IntersectK through SvS (svs.rs:50-73), sets in the order given.
    template <class V>
    void kset(const std::vector<Slice> &sets, V &visitor) const
    {
        std::vector<const uint32_t *> ptrs;
        std::vector<size_t> lens;
        size_t cap = 1;
        for (const Slice &s : sets) {
            ptrs.push_back(s.ptr);
            lens.push_back(s.len);
        }
        if (!sets.empty()) cap = std::max<size_t>(1, std::min_element(lens.begin(), lens.end())[0]);
        std::vector<uint32_t> buf(cap);
        size_t n = 0;
        check(setops_hip_svs(id_, ptrs.data(), lens.data(), sets.size(), buf.data(), buf.size(), &n));
        for (size_t i = 0; i < n; i++) visitor.visit(buf[i]);
    }

private:
    int id_ = -1;
    std::string name_;
    unsigned flags_ = 0;
};

inline std::vector<std::string> registered()
{
    std::vector<std::string> v;
    for (int i = 0; i < setops_hip_algo_count(); i++) v.emplace_back(setops_hip_algo_name(i));
    return v;
}

inline void init(int device = 0) { check(setops_hip_init(device)); }

// ---- runners (intersect.rs:44-90) -------------------------------------------------------------------

inline std::vector<uint32_t> run_2set(Slice a, Slice b, const Entry &f)
{
    VecWriter w;
    f(a, b, w);
    return w.into();
}

inline std::vector<uint32_t> run_2set_c(Slice a, Slice b, const Entry &f)
{
    std::vector<uint32_t> r(std::min(a.len, b.len));
    r.resize(f.intersect_c(a, b, r.data(), r.size()));
    return r;
}

inline std::vector<uint32_t> run_kset(const std::vector<Slice> &sets, const Entry &f)
{
    if (sets.size() < 2) throw Error(SETOPS_ERR_INVALID, "run_kset needs at least 2 sets");
    VecWriter w;
    f.kset(sets, w);
    return w.into();
}

inline BsrVec run_2set_bsr(BsrRef a, BsrRef b, const Entry &f)
{
    BsrVec w;
    f.bsr(a, b, w);
    return w;
}

// ---- SvS (svs.rs) -----------------------------------------------------------------------------------

// sets[0] ∩ sets[1] into `left`, then (swap; clear; right ∩ next set into left) for the rest;
// returns the visitor that holds the result.  V needs visit/clear/as_ref.
template <class V>
V &svs_generic(const std::vector<Slice> &sets, V &left, V &right, const Entry &f)
{
    if (sets.size() < 2) throw Error(SETOPS_ERR_INVALID, "svs_generic needs at least 2 sets (svs.rs:61)");
    V *l = &left, *r = &right;
    f(sets[0], sets[1], *l);
    for (size_t i = 2; i < sets.size(); i++) {
        std::swap(l, r);
        l->clear();
        const auto prev = r->as_ref();
        const std::vector<uint32_t> cur(prev.begin(), prev.end());
        f(Slice(cur), sets[i], *l);
    }
    return *l;
}

inline std::vector<uint32_t> run_svs(const std::vector<Slice> &sets, const Entry &f)
{
    VecWriter left, right;
    return svs_generic(sets, left, right, f).into();
}

// ---- timing (harness.rs:35-62) ----------------------------------------------------------------------

class Harness {
public:
    explicit Harness(std::chrono::duration<double> warmup) : warmup_(warmup) {}
    // Runs prepare()+run(data) until the warmup has elapsed, then one timed run; returns the
    // timed run's seconds and its data.
    template <class Prepare, class Run>
    auto time(Prepare prepare, Run run) -> std::pair<double, decltype(prepare())>
    {
        using clock = std::chrono::steady_clock;
        const auto w0 = clock::now();
        while (clock::now() - w0 < warmup_) {
            auto d = prepare();
            run(d);
        }
        auto data = prepare();
        const auto t0 = clock::now();
        run(data);
        const double s = std::chrono::duration<double>(clock::now() - t0).count();
        return {s, std::move(data)};
    }

private:
    std::chrono::duration<double> warmup_;
};

}  // namespace setops
