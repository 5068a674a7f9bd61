// Entries of the carry-sweep kernel tables (CarryKernel, ll_kernels.h) for
// the ll_carry_*.hip instantiation files.
#pragma once
#include "ll_sweeps.h"

// carry_sweep<W, R, T> with its load / unload kernels and line table
#define LL_CARRY_T(W, R, T)                                                              \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr}
#define LL_CARRY(W, R) LL_CARRY_T(W, R, 256)

// with the persistent form, for schedules whose blocks take several rounds
#define LL_CARRY_PIPE(W, R, T)                                                           \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr,       \
   (const void *)carry_sweep_pipe<W, R, T>}

// 256 threads, with the 512-thread form for schedules whose blocks fit in
// one round of it
#define LL_CARRY_W(W, R)                                                                 \
  {W, R, (const void *)carry_sweep<W, R, 256>, (const void *)load_lines<W, R, 256>,      \
   (const void *)unload_lines<W, R, 256>, 256,                                           \
   carry_pass_table<W, R, 256>() ? (const void *)line_pass_table<W, R, 256> : nullptr,   \
   nullptr, carry_wide_form<W, R>()}

// 256 threads, with the persistent 512-thread form for schedules whose
// blocks take several rounds
#define LL_CARRY_HP(W, R)                                                                \
  {W, R, (const void *)carry_sweep<W, R, 256>, (const void *)load_lines<W, R, 256>,      \
   (const void *)unload_lines<W, R, 256>, 256,                                           \
   carry_pass_table<W, R, 256>() ? (const void *)line_pass_table<W, R, 256> : nullptr,   \
   carry_pipe_form<W, R>(), nullptr}
