/* Greedy bit-mask colouring of iteration elements -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the algorithm documented in the reference at
 * doc/sphinx/source/plan.rst:46-80 (the plan module it describes is not in
 * this version of the reference):
 *   for each element, in element order:
 *     1. OR together the colour masks of every set element it references
 *        through the conflicting (INC/RW/WRITE) maps;
 *     2. take the lowest free bit as its colour;
 *     3. set that bit in the masks of every referenced set element.
 *   Masks are 32 bits; elements that find no free bit wait for another pass,
 *   which starts with cleared masks and adds 32 to the colours it hands out.
 * Map entries below zero reference nothing.  Maps that point into the same
 * target set share one mask array (target_id).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int32_t op2o_greedy_colour(int32_t nelem, int32_t nmaps,
                           const int32_t *const *maps, const int32_t *arity,
                           const int32_t *target_id, const int32_t *target_size,
                           int32_t ntargets, int32_t *colour)
{
  uint32_t **masks = (uint32_t **)calloc((size_t)ntargets, sizeof(uint32_t *));
  for (int32_t t = 0; t < ntargets; ++t)
    masks[t] = (uint32_t *)calloc((size_t)(target_size[t] > 0 ? target_size[t] : 1), sizeof(uint32_t));
  for (int32_t e = 0; e < nelem; ++e) colour[e] = -1;
  int32_t done = 0, offset = 0, ncolours = 0;
  while (done < nelem) {
    for (int32_t t = 0; t < ntargets; ++t)
      memset(masks[t], 0, sizeof(uint32_t) * (size_t)(target_size[t] > 0 ? target_size[t] : 1));
    for (int32_t e = 0; e < nelem; ++e) {
      if (colour[e] >= 0) continue;
      uint32_t mask = 0;
      for (int32_t m = 0; m < nmaps; ++m)
        for (int32_t i = 0; i < arity[m]; ++i) {
          const int32_t x = maps[m][(int64_t)e * arity[m] + i];
          if (x >= 0) mask |= masks[target_id[m]][x];
        }
      if (mask == 0xFFFFFFFFu) continue;
      int32_t c = 0;
      while (mask & (1u << c)) ++c;
      colour[e] = offset + c;
      if (offset + c + 1 > ncolours) ncolours = offset + c + 1;
      ++done;
      for (int32_t m = 0; m < nmaps; ++m)
        for (int32_t i = 0; i < arity[m]; ++i) {
          const int32_t x = maps[m][(int64_t)e * arity[m] + i];
          if (x >= 0) masks[target_id[m]][x] |= (1u << c);
        }
    }
    offset += 32;
  }
  for (int32_t t = 0; t < ntargets; ++t) free(masks[t]);
  free(masks);
  return ncolours;
}
