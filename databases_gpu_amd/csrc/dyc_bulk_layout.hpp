// dyc_bulk_layout.hpp -- host side of the group launches (internal to libdyc): how one batch of a
// group (keys grouped by table by host offsets) is laid out for the kernels of dyc_bulk.hip, and
// the find's work items by XCD.
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

// Find items laid out by XCD. Workgroups go to the device's XCDs in turn (workgroup b to XCD
// b % kXcds) and a probe launch's wave w of workgroup b takes item slots b * 4 + w, then every
// (grid * 4)-th slot after it, so slot i runs on XCD (i / 4) % kXcds (for a grid that is a
// multiple of kXcds; any other grid only loses the layout's benefit). Each table goes whole to
// one XCD (largest first, to the one with the fewest items), so the waves of an XCD probe a few
// tables at a time and its L2 holds their lines, instead of every XCD reading every table the
// launch has in flight. Slots an XCD has no item for hold empty items.
constexpr uint32_t kXcds = 8, kProbeWavesPerBlock = 4;
constexpr uint64_t kXcdLayoutKeys = 1ull << 22;  // eager finds of fewer keys keep the table order

// The layout itself: n tables' batches (offsets, as the calls take them) cut into items of at
// most per keys. The caller checked that each table's batch fits a 32-bit length.
inline std::vector<GroupItem> items_by_xcd(const uint64_t *offsets, uint32_t n, uint64_t per)
{
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < n; ++i)
        if (offsets[i + 1] > offsets[i]) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return offsets[a + 1] - offsets[a] > offsets[b + 1] - offsets[b];
    });
    std::vector<GroupItem> seq[kXcds];
    for (uint32_t t : order) {
        uint32_t x = 0;
        for (uint32_t y = 1; y < kXcds; ++y)
            if (seq[y].size() < seq[x].size()) x = y;
        for (uint64_t b = offsets[t]; b < offsets[t + 1]; b += per)
            seq[x].push_back(GroupItem{b, (uint32_t)std::min<uint64_t>(per, offsets[t + 1] - b), t});
    }
    size_t most = 0;
    for (const auto &q : seq) most = std::max(most, q.size());
    const uint64_t rows = (most + kProbeWavesPerBlock - 1) / kProbeWavesPerBlock;  // slot groups per XCD
    std::vector<GroupItem> items((size_t)(rows * kProbeWavesPerBlock * kXcds), GroupItem{0, 0, 0});
    for (size_t i = 0; i < items.size(); ++i) {
        const size_t x = (i / kProbeWavesPerBlock) % kXcds;
        const size_t j = (i / (kProbeWavesPerBlock * kXcds)) * kProbeWavesPerBlock + i % kProbeWavesPerBlock;
        if (j < seq[x].size()) items[i] = seq[x][j];
    }
    return items;
}

}  // namespace host
}  // namespace dyc
