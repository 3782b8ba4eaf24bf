// server.hip -- request loop (filled in below).
#include "engine.h"

extern "C" {
int ccd_server_start(const ccd_server_config*, ccd_server_handle*) { return CCD_ERR_UNSUPPORTED; }
int ccd_server_submit(ccd_server_handle, const ccd_data_desc*, const ccd_fit_args*, const double*, int64_t*) {
    return CCD_ERR_UNSUPPORTED;
}
int ccd_server_poll(ccd_server_handle, int64_t, int, double*, ccd_fit_result*, ccd_request_stats*) {
    return CCD_ERR_UNSUPPORTED;
}
int ccd_server_stop(ccd_server_handle) { return CCD_ERR_UNSUPPORTED; }
}
