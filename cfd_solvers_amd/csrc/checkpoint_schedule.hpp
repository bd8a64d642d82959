// checkpoint_schedule.hpp -- binomial checkpointing for reversing a chain of time steps.
//
// Reversing steps [a, b) means running the transposed step at t = b-1, ..., a, each of which needs
// the forward state f_t.  State a is held in a buffer; s further state buffers are free.  The plan
// advances a -> m into one free buffer, reverses [m, b) with s - 1 buffers, releases the buffer and
// reverses [a, m) with s buffers (the left part again keeps its own start state).  With
//
//     beta(s, r) = C(s + r, s)
//
// a segment of length l <= beta(s, r) is reversed with every step recomputed at most r times, and
// the minimal number of forward steps (the first sweep included) is
//
//     F(l, s) = r l - beta(s + 1, r - 1),   beta(s, r - 1) < l <= beta(s, r)
//
// (Griewank, "Achieving logarithmic growth of temporal and spatial complexity in reverse automatic
// differentiation", Optim. Methods Softw. 1, 1992; Griewank & Walther, ACM TOMS 26, 2000).
// F(., s) is piecewise linear with slope r, so the cost m + F(l - m, s - 1) + F(m, s) of a split
// is flat exactly while the left part runs at repetition r - 1 and the right part at r:
//
//     max(beta(s, r - 2), l - beta(s - 1, r)) <= m - a <= min(beta(s, r - 1), l - beta(s - 1, r - 1))
//
// and every split in that range attains F(l, s); split() takes the lower end.  Pure host code; no
// device calls.
#pragma once

#include <cstdint>
#include <limits>
#include <stdexcept>

namespace cfd_hip {

struct BinomialSchedule {
    static constexpr uint64_t kSaturated = std::numeric_limits<uint64_t>::max();

    // C(s + r, s), saturating at kSaturated.
    static uint64_t beta(uint64_t s, uint64_t r) {
        uint64_t v = 1;
        const uint64_t k = s < r ? s : r;
        for (uint64_t i = 1; i <= k; ++i) {
            const uint64_t num = s + r - k + i;  // v * num / i stays integral: v = C(s + r - k + i - 1, i - 1)
            if (v > kSaturated / num) return kSaturated;
            v = v * num / i;
        }
        return v;
    }

    // Smallest r with beta(s, r) >= l (l >= 1, s >= 1).
    static uint64_t repetitions(uint64_t l, uint64_t s) {
        if (s == 0) throw std::invalid_argument("binomial checkpointing needs at least one free state buffer");
        uint64_t r = 0;
        while (beta(s, r) < l) ++r;
        return r;
    }

    // Length of the left part when [a, a + l) is split with s free buffers (l >= 2, s >= 1).
    static uint64_t split(uint64_t l, uint64_t s) {
        const uint64_t r = repetitions(l, s);
        uint64_t m = r >= 2 ? beta(s, r - 2) : 1;
        const uint64_t right = beta(s - 1, r);  // the right part may hold at most this many steps
        if (l > right && l - right > m) m = l - right;
        return m < l ? m : l - 1;
    }

    // Forward steps the plan executes for a segment of length l with s free buffers (closed form).
    static uint64_t forward_steps(uint64_t l, uint64_t s) {
        if (l <= 1) return 0;
        const uint64_t r = repetitions(l, s);
        return r * l - beta(s + 1, r - 1);
    }

    // Forward steps the plan executes, counted by walking it (equals forward_steps; used by tests).
    static uint64_t walk(uint64_t l, uint64_t s) {
        uint64_t steps = 0;
        while (l > 1) {
            const uint64_t m = split(l, s);
            steps += m + walk(l - m, s - 1);
            l = m;
        }
        return steps;
    }
};

}  // namespace cfd_hip
