#pragma once

#include <helix/examples.h>
#include <helix/fwd.h>
#include <helix/io.h>
#include <helix/types.h>
#include <helix/version.h>
