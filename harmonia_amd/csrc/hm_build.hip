// Device-side bulk build of the Harmonia layout from sorted keys.
//
// The reference builds a pointer tree by repeated inserts and then re-lays it out breadth-first
// with a prefix-sum child array on the host (BPlusTree::sort / prepareArray,
// H/src/traversal.cpp:35-195; generateCFB / traverse_transform, H/regular_bpt/traversal.cpp:36-99).
// Here the sorted key set is cut into leaves of `fill` keys, and every inner level is emitted
// bottom-up in parallel: separator s of node j is the minimum key of child j*F + s + 1 (the
// separator = first key of the right sibling, H/src/bpt.cpp:36), unused slots hold KEY_MAX
// (H/regular_bpt/traversal.cpp:52-54), and the index granule holds key[8j+7]
// (Harmonia's inner_index, H/src/bpt.h:76-81).  The prefix array is then one device scan.
#include "hm_internal.h"

namespace hm {
namespace {

// Both reductions below end in atomics on one address, which the memory system serializes (~10 ns
// each): every block reduces in LDS first and the grid is capped (kReduceBlocks), so a launch issues
// at most a few thousand of them.
constexpr int kReduceBlocks = 4096;
constexpr int kUnroll = 4;  // independent loads in flight per thread

__global__ __launch_bounds__(256) void validate_sorted_kernel(const uint64_t* __restrict__ k, size_t n,
                                                              uint32_t* __restrict__ flags) {
  __shared__ uint32_t block_f;
  if (threadIdx.x == 0) block_f = 0;
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  uint32_t f = 0;
  for (size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += kUnroll * stride) {
    uint64_t x[kUnroll], y[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = b + u * stride;
      x[u] = i < n ? k[i] : 0ull;
      y[u] = (i < n && i > 0) ? k[i - 1] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = b + u * stride;
      if (i < n) {
        if (x[u] == KEY_MAX) f |= 1u;
        if (i > 0 && y[u] >= x[u]) f |= 2u;
      }
    }
  }
  const uint64_t b1 = __ballot((f & 1u) != 0), b2 = __ballot((f & 2u) != 0);
  const uint32_t w = (b1 ? 1u : 0u) | (b2 ? 2u : 0u);
  if ((threadIdx.x & 63) == 0 && w) atomicOr(&block_f, w);
  __syncthreads();
  if (threadIdx.x == 0 && block_f) atomicOr(flags, block_f);
}

__global__ __launch_bounds__(256) void key_bits_kernel(const uint64_t* __restrict__ k, size_t n,
                                                       uint64_t* __restrict__ orand) {
  __shared__ unsigned long long block_o, block_a;
  if (threadIdx.x == 0) {
    block_o = 0ull;
    block_a = ~0ull;
  }
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  uint64_t o = 0, a = ~0ull;
  for (size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += kUnroll * stride) {
    uint64_t x[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const size_t i = b + u * stride;
      x[u] = i < n ? k[i] : k[b];  // k[b] leaves the or / and unchanged
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      o |= x[u];
      a &= x[u];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    o |= __shfl_xor(o, off, 64);
    a &= __shfl_xor(a, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicOr(&block_o, (unsigned long long)o);
    atomicAnd(&block_a, (unsigned long long)a);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicOr((unsigned long long*)&orand[0], block_o);
    atomicAnd((unsigned long long*)&orand[1], block_a);
  }
}

__global__ void fill_leaves_kernel(const uint64_t* __restrict__ keys, const uint64_t* __restrict__ vals, size_t n,
                                   int fill, Leaf* __restrict__ leaves, size_t n_leaves) {
  const size_t total = n_leaves * LEAF_KEYS;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const size_t leaf = t / LEAF_KEYS;
    const int slot = (int)(t % LEAF_KEYS);
    const size_t src = leaf * (size_t)fill + slot;
    const bool have = slot < fill && src < n;
    leaves[leaf].key[slot] = have ? keys[src] : KEY_MAX;
    leaves[leaf].val[slot] = have ? vals[src] : 0ull;
  }
}

__global__ void leaf_min_kernel(const Leaf* __restrict__ leaves, size_t n_leaves, uint64_t* __restrict__ mins) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_leaves; i += stride)
    mins[i] = leaves[i].key[0];
}

// One thread per node slot (NS slots per node).
template <int F>
__global__ void build_level_kernel(const uint64_t* __restrict__ child_min, size_t m_child,
                                   uint64_t* __restrict__ nodes, uint64_t* __restrict__ node_min,
                                   uint32_t* __restrict__ nchild, size_t m) {
  constexpr int IS = idx_slots(F);
  constexpr int NS = node_slots(F);
  const size_t total = m * NS;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const size_t j = t / NS;
    const int s = (int)(t % NS);
    const size_t first = j * (size_t)F;
    int key_slot;
    if (s < IS) {
      key_slot = (s < F / 8) ? 8 * s + 7 : F - 1;  // F-1 is always padding
    } else {
      key_slot = s - IS;
    }
    const size_t c = first + (size_t)key_slot + 1;  // separator key_slot = min of child key_slot + 1
    const uint64_t v = (key_slot < F - 1 && c < m_child) ? child_min[c] : KEY_MAX;
    nodes[t] = v;
    if (s == 0) {
      node_min[j] = child_min[first];
      const size_t rem = m_child - first;
      nchild[j] = (uint32_t)(rem < (size_t)F ? rem : (size_t)F);
    }
  }
}

// prefix[node] into the node's last index slot (canonically KEY_MAX), one thread per node.
template <int F>
__global__ void fold_prefix_kernel(uint64_t* __restrict__ inner, const uint32_t* __restrict__ prefix, size_t n) {
  constexpr int IS = idx_slots(F);
  constexpr int NS = node_slots(F);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
    inner[j * NS + IS - 1] = (uint64_t)prefix[j];
}

int grid_for(size_t work, int threads) {
  size_t g = (work + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

}  // namespace

hipError_t validate_sorted(const uint64_t* keys, size_t n, uint32_t* d_flags, hipStream_t s) {
  if (n == 0) return hipSuccess;
  ProfScope prof("build", s);
  int g = grid_for(n, 256);
  if (g > kReduceBlocks) g = kReduceBlocks;
  hipLaunchKernelGGL(validate_sorted_kernel, dim3(g), dim3(256), 0, s, keys, n, d_flags);
  return hipGetLastError();
}

hipError_t key_bit_summary(const uint64_t* keys, size_t n, uint64_t* d_orand, hipStream_t s) {
  if (n == 0) return hipSuccess;
  ProfScope prof("build", s);
  int g = grid_for(n, 256);
  if (g > kReduceBlocks) g = kReduceBlocks;
  hipLaunchKernelGGL(key_bits_kernel, dim3(g), dim3(256), 0, s, keys, n, d_orand);
  return hipGetLastError();
}

hipError_t fill_leaves(const uint64_t* keys, const uint64_t* vals, size_t n, int fill, Leaf* leaves, size_t n_leaves,
                       hipStream_t s) {
  ProfScope prof("build", s);
  hipLaunchKernelGGL(fill_leaves_kernel, dim3(grid_for(n_leaves * LEAF_KEYS, 256)), dim3(256), 0, s, keys, vals, n,
                     fill, leaves, n_leaves);
  return hipGetLastError();
}

hipError_t leaf_min_keys(const Leaf* leaves, size_t n_leaves, uint64_t* child_min, hipStream_t s) {
  ProfScope prof("build", s);
  hipLaunchKernelGGL(leaf_min_kernel, dim3(grid_for(n_leaves, 256)), dim3(256), 0, s, leaves, n_leaves, child_min);
  return hipGetLastError();
}

hipError_t build_level(const uint64_t* child_min, size_t m_child, int F, uint64_t* nodes, uint64_t* node_min,
                       uint32_t* nchild, size_t m, hipStream_t s) {
  ProfScope prof("build", s);
  const int g = grid_for(m * (size_t)node_slots(F), 256);
  switch (F) {
    case 32:
      hipLaunchKernelGGL(build_level_kernel<32>, dim3(g), dim3(256), 0, s, child_min, m_child, nodes, node_min, nchild,
                         m);
      break;
    case 64:
      hipLaunchKernelGGL(build_level_kernel<64>, dim3(g), dim3(256), 0, s, child_min, m_child, nodes, node_min, nchild,
                         m);
      break;
    case 128:
      hipLaunchKernelGGL(build_level_kernel<128>, dim3(g), dim3(256), 0, s, child_min, m_child, nodes, node_min,
                         nchild, m);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t fold_prefix(uint64_t* inner, const uint32_t* prefix, size_t n_inner, int F, hipStream_t s) {
  if (n_inner == 0) return hipSuccess;
  ProfScope prof("build", s);
  const int g = grid_for(n_inner, 256);
  switch (F) {
    case 32: hipLaunchKernelGGL(fold_prefix_kernel<32>, dim3(g), dim3(256), 0, s, inner, prefix, n_inner); break;
    case 64: hipLaunchKernelGGL(fold_prefix_kernel<64>, dim3(g), dim3(256), 0, s, inner, prefix, n_inner); break;
    case 128: hipLaunchKernelGGL(fold_prefix_kernel<128>, dim3(g), dim3(256), 0, s, inner, prefix, n_inner); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace hm
