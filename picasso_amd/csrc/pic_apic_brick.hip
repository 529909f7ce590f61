// APIC brick scheme: particles binned by brick, LDS node accumulation.
#include "pic_apic.hpp"

namespace pic {

int apic_p2g_brick(pic_context*, int, int, int, const pic_mesh&, int64_t, const pic_pfield&,
                   const pic_pfield&, const pic_pfield&, const pic_grid&, const pic_grid&,
                   const pic_bins&)
{
    return PIC_ERR_UNSUPPORTED;
}

int apic_g2p_brick(pic_context*, int, int, int, const pic_mesh&, int64_t, const pic_pfield&,
                   const pic_grid&, const pic_pfield&, const pic_bins&)
{
    return PIC_ERR_UNSUPPORTED;
}

}  // namespace pic
