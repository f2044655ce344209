/*
 * kp_compact_oracle.c -- CPU restatement of the packet-pool compaction and
 * the per-destination binning, TEST INFRASTRUCTURE ONLY.
 *
 * or_pool_keep follows pool_t::keep_pool (src/modules/particle/types/
 * particle_pool.h:191-248): packets marked for removal leave holes; the
 * n_keep = n - n_removed survivors end up in [0, n_keep) by moving kept
 * packets from the tail [n_keep, n) into the holes of the head, nothing else
 * moves.  The reference pairs hole map_pars[i] (an atomic_inc list, so in
 * thread-completion order) with tail packet n-1-i; the pairing here is the
 * deterministic one: the j-th hole of the head (ascending index) receives
 * the j-th kept packet of the tail counted from the end.
 *
 * or_bin_by_key follows comm_t::reg / cpy_d (src/modules/particle/boundary/
 * comm.h:118-127, 226-239): every packet with a destination d in
 * [0, n_bins) gets a slot in destination d's buffer, slot = offsets[d] +
 * sequence number.  The reference draws the sequence number with
 * atomicAdd(n_send[d]) (completion order); here it is the packet's rank
 * among the packets of d in index order (a stable counting sort).  Keys
 * outside [0, n_bins) stay (the reference's "no send").
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "kp_oracle.h"

int64_t or_pool_keep(void *pool, int64_t n, int64_t unit, const uint8_t *keep)
{
    unsigned char *p = (unsigned char *)pool;
    int64_t n_keep = 0;
    for (int64_t i = 0; i < n; ++i)
        n_keep += keep[i] != 0;
    int64_t h = 0, t = n - 1;
    for (;;) {
        while (h < n_keep && keep[h])
            ++h;
        while (t >= n_keep && !keep[t])
            --t;
        if (h >= n_keep || t < n_keep)
            break;
        memcpy(p + h * unit, p + t * unit, (size_t)unit);
        ++h;
        --t;
    }
    return n_keep;
}

int64_t or_bin_by_key(const int32_t *key, int64_t n, int32_t n_bins, int64_t *offsets,
                      int32_t *order, int32_t *slot)
{
    memset(offsets, 0, sizeof(int64_t) * ((size_t)n_bins + 1));
    for (int64_t i = 0; i < n; ++i)
        if (key[i] >= 0 && key[i] < n_bins)
            offsets[key[i] + 1] += 1;
    for (int32_t b = 0; b < n_bins; ++b)
        offsets[b + 1] += offsets[b];
    int64_t *next = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n_bins + 1));
    memcpy(next, offsets, sizeof(int64_t) * ((size_t)n_bins + 1));
    for (int64_t i = 0; i < n; ++i) {
        if (key[i] >= 0 && key[i] < n_bins) {
            const int64_t s = next[key[i]]++;
            order[s] = (int32_t)i;
            if (slot)
                slot[i] = (int32_t)s;
        } else if (slot) {
            slot[i] = -1;
        }
    }
    free(next);
    return offsets[n_bins];
}

void or_gather(const void *src, int64_t unit, const int32_t *order, int64_t n, void *dst)
{
    for (int64_t s = 0; s < n; ++s)
        memcpy((unsigned char *)dst + s * unit, (const unsigned char *)src + order[s] * unit,
               (size_t)unit);
}
