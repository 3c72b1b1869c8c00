/* Stack clearing helper of the reference harness (test infrastructure):
 * zeroes `bytes` of stack below the caller's frame, so the automatic arrays
 * of the next call at the same depth start from zero. */
#include <stddef.h>
#include <string.h>

__attribute__((noinline)) void ccs_ref_clear_stack(size_t bytes) {
  char* buf = __builtin_alloca(bytes);
  memset(buf, 0, bytes);
  __asm__ volatile("" : : "r"(buf) : "memory");
}
