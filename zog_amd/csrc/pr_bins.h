// PageRank row binning by predecessor count (rows = live vertices).
#pragma once
#define ZG_PR_SMALL 16     // <= 16 predecessors: one lane per row
#define ZG_PR_MED 2048     // <= 2048: one wavefront per row
#define ZG_PR_SEG 4096     // longer rows: 256-lane blocks over 4096-arc segments
#define ZG_PR_BLOCK 256
