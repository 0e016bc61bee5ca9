// dyc_bulk_layout.hpp -- host side of the group bulk fill (internal to libdyc): how one batch of a
// group (keys grouped by table by host offsets) is laid out for the kernels of dyc_bulk.hip.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "dyc_kernels.hpp"

namespace dyc {
namespace host {

// The participating tables, their partition blocks (at most kGroupBulkBlockKeys keys of one
// table each) and each block's table index.
struct BulkLayout {
    std::vector<BulkTable> tabs;
    std::vector<GroupItem> blocks;
    std::vector<uint32_t> btab;
    uint64_t total = 0;
    uint32_t chunks = 0;
};

// Lays out a batch of n_tables tables (desc(i): table i's Desc) for the bulk fill; false when it
// does not qualify: a participating table has more than kGroupBulkChunks chunks, or the batch
// has fewer than two keys per bucket of those tables, or 2^32 keys or more.
template <typename DescOf>
bool bulk_layout(uint32_t n_tables, const uint64_t *offsets, DescOf desc, BulkLayout *L)
{
    L->total = offsets[n_tables];
    if (L->total == 0 || L->total >= (1ull << 32)) return false;
    uint64_t buckets = 0;
    for (uint32_t i = 0; i < n_tables; ++i) {
        const uint64_t n = offsets[i + 1] - offsets[i];
        if (!n) continue;
        const Desc &D = desc(i);
        uint32_t ch = 0;
        for (uint32_t t = 0; t < D.d; ++t) {
            ch += (D.nb[t] + kChunkBuckets - 1) / kChunkBuckets;
            buckets += D.nb[t];
        }
        if (ch > kGroupBulkChunks) return false;
        const uint32_t j = (uint32_t)L->tabs.size();
        BulkTable T{offsets[i], (uint32_t)n, i, (uint32_t)L->blocks.size(), 0, L->chunks, ch};
        for (uint64_t b = offsets[i]; b < offsets[i + 1]; b += kGroupBulkBlockKeys) {
            L->blocks.push_back(GroupItem{b, (uint32_t)std::min<uint64_t>(kGroupBulkBlockKeys, offsets[i + 1] - b), i});
            L->btab.push_back(j);
        }
        T.blocks = (uint32_t)L->blocks.size() - T.first_block;
        L->chunks += ch;
        L->tabs.push_back(T);
    }
    return L->total >= 2 * buckets;
}

}  // namespace host
}  // namespace dyc
