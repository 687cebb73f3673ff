// PageRank row binning by predecessor count (rows = live vertices).  A row
// is summed by a group of L lanes (one wavefront holds 64/L rows) so short
// rows do not leave most of a wavefront idle.
#pragma once
#ifndef ZG_PR_SMALL
#define ZG_PR_SMALL 4      // <= 4 predecessors: one lane per row
#endif
#ifndef ZG_PR_M4
#define ZG_PR_M4 32        // <= 32: 4 lanes per row
#endif
#ifndef ZG_PR_M16
#define ZG_PR_M16 256      // <= 256: 16 lanes per row
#endif
#define ZG_PR_MED 2048     // <= 2048: 64 lanes per row
#define ZG_PR_SEG 4096     // longer rows: listed as 4096-arc segments (zg_graph_table), summed by a whole block
#define ZG_PR_BLOCK 256
#ifndef ZG_PR_XCACHE_BYTES
#define ZG_PR_XCACHE_BYTES 16384  // LDS copy of the most-gathered values per 256-thread iteration block
#endif
