// CPU stand-in for the part of the HIP device API that the solver kernels use,
// so that openusct_amd/csrc/usct_kernels.hpp compiles with g++ and its kernels
// run lane by lane on the host (test infrastructure only; never linked into the
// library).  A wavefront's 64 lanes run as 64 cooperative contexts on the
// calling thread (x86-64 only: a six-register stack switch, cheap next to
// ucontext, whose signal-mask syscall and sanitizer hook dominate otherwise),
// switched round-robin, one work-group at a time: __shfl_xor stores the lane's
// value, yields until every live lane has stored, reads its partner's, and
// yields again before the slot can be reused; __syncthreads yields once.  A
// lane that returns early drops out, as an inactive lane would.  __shared__
// arrays are static: one instance, shared by the lanes of the running
// work-group (work-groups run one after another).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __builtin_amdgcn_readfirstlane(v) (v)

struct float2 {
    float x, y;
};
inline float2 make_float2(float x, float y) { return {x, y}; }

inline float __uint_as_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint32_t __float_as_uint(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }


namespace emu {

struct U3 {
    unsigned x = 0, y = 0, z = 0;
};
inline U3 tid, bid, gdim, bdim;

// switch(save, to): push the callee-saved registers, store the stack pointer in
// *save, load `to`, pop the registers and return into the other context.
extern "C" void emu_switch(void **save, void *to);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch, @function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

constexpr size_t kStack = 256 << 10;
constexpr int kMaxLanes = 256;  // one work-group

struct Wave {
    void *sched = nullptr;
    void *lane[kMaxLanes] = {};
    std::vector<std::unique_ptr<char[]>> stacks;
    bool done[kMaxLanes] = {};
    int cur = 0, nlanes = 0;
    unsigned block = 0;
    const std::function<void()> *fn = nullptr;
    float xch[kMaxLanes] = {};
};
inline Wave W;

inline void enter(unsigned l) {
    tid.x = l;
    bid.x = W.block;
}

[[noreturn]] inline void lane_main() {
    const int l = W.cur;
    enter((unsigned)l);
    (*W.fn)();
    W.done[l] = true;
    emu_switch(&W.lane[l], W.sched);
    __builtin_unreachable();
}

// Back to the scheduler; resumes after every other live lane of the
// work-group has yielded once (a barrier over the live lanes).
inline void yield() {
    const int l = W.cur;
    emu_switch(&W.lane[l], W.sched);
    enter((unsigned)l);
}

// All lanes of one work-group run together, round-robin; a round ends when
// every live lane has yielded, so __syncthreads and the shuffle exchanges are
// barriers over the work-group (a wave's lanes yield in lockstep).
inline void run_block(unsigned block, unsigned threads, const std::function<void()> &fn) {
    if (W.stacks.empty())
        for (int l = 0; l < kMaxLanes; ++l) W.stacks.emplace_back(new char[kStack]);
    W.block = block;
    W.fn = &fn;
    W.nlanes = (int)threads;
    for (int l = 0; l < W.nlanes; ++l) {
        W.done[l] = false;
        // initial frame: six zero registers, the entry address, a zero return
        // address slot; the entry sees the stack pointer as after a call
        uintptr_t top = (reinterpret_cast<uintptr_t>(W.stacks[l].get()) + kStack) & ~uintptr_t(15);
        void **sp = reinterpret_cast<void **>(top);
        *--sp = nullptr;
        *--sp = reinterpret_cast<void *>(&lane_main);
        for (int r = 0; r < 6; ++r) *--sp = nullptr;
        W.lane[l] = sp;
    }
    for (bool live = true; live;) {
        live = false;
        for (int l = 0; l < W.nlanes; ++l) {
            if (W.done[l]) continue;
            W.cur = l;
            emu_switch(&W.sched, W.lane[l]);
            live |= !W.done[l];
        }
    }
}

// Run a kernel body over `blocks` work-groups of `threads` lanes (<= 256).
inline void launch(unsigned blocks, unsigned threads, const std::function<void()> &body) {
    gdim = U3{blocks, 1, 1};
    bdim = U3{threads, 1, 1};
    for (unsigned b = 0; b < blocks; ++b) run_block(b, threads, body);
}

inline float shfl_xor(float v, int off) {
    const int l = W.cur;  // work-group lane; the partner is in the same wavefront
    W.xch[l] = v;
    yield();
    const float r = W.xch[(l & ~63) | ((l & 63) ^ off)];
    yield();
    return r;
}

}  // namespace emu

#define threadIdx emu::tid
#define blockIdx emu::bid
#define gridDim emu::gdim
#define blockDim emu::bdim

inline float __shfl_xor(float v, int off, int /*width*/) { return emu::shfl_xor(v, off); }
inline void __syncthreads() { emu::yield(); }
