from xlb_amd.operator.stream.stream import Stream as Stream
