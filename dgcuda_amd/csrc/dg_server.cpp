#include "dg_server.h"

namespace dgmi {

Status Server::create(const dg_server_config&, Server**) { return Status::err(DG_ERR_UNSUPPORTED, "server not built yet"); }

}  // namespace dgmi
