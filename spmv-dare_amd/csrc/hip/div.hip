// DIV-RF on gfx950 (placeholder until the codec lands).
#include "format.hpp"
namespace spmvhip {
int div_convert(spmvhip_format *, const int32_t *, const int32_t *, const double *)
{
	return set_error(SPMVHIP_EINVAL, "div_rf: not available in this build");
}
int div_spmv(spmvhip_format *, const double *, double *, hipStream_t)
{
	return set_error(SPMVHIP_EINVAL, "div_rf: not available in this build");
}
int div_export_decoded(const spmvhip_format *, int, void *, size_t *)
{
	return set_error(SPMVHIP_EINVAL, "div_rf: not available in this build");
}
} // namespace spmvhip
