// gg_internal.h -- declarations shared by the evaluator's translation units.
#ifndef GG_INTERNAL_H
#define GG_INTERNAL_H
namespace gg {
int set_error(int code, const char *fmt, ...);
const char *last_error();
int last_error_code();
}  // namespace gg
#endif
