"""HTTP transport for the GPU page-query endpoints.

Serves POST /query/page/gpu and POST /query/page/gpu/batch (the routes of
crates/morphogen-server/src/network/api/mod.rs:886-896) from a thread pool;
every request is handed unchanged to the C ABI handler (mg_wire_handle via
WireApi.handle), which parses, validates, runs the keys through the GPU
request loop and formats the JSON answer. This module only moves bytes.

CLI (a seeded synthetic matrix; pages are rows of --row-bytes):
    python -m morphogenesis_amd.http_api --rows-log2 12 --row-bytes 4096 --port 8080
"""
import argparse
import sys
import threading
from http.server import BaseHTTPRequestHandler, ThreadingHTTPServer

from . import WIRE_MAX_BODY, GpuPageMatrix, GpuScanner, GpuServer, WireApi


def make_handler(api: WireApi):
    limit = api.opts.max_body_bytes or WIRE_MAX_BODY

    class Handler(BaseHTTPRequestHandler):
        protocol_version = "HTTP/1.1"

        def _serve(self):
            n = int(self.headers.get("Content-Length") or 0)
            # Read at most one byte past the limit: the handler answers 413.
            body = self.rfile.read(min(n, limit + 1)) if n > 0 else b""
            if n > limit:
                self.close_connection = True
            status, out, err = api.handle(self.command, self.path.split("?", 1)[0],
                                          self.headers.get("Content-Type"), body)
            self.send_response(status)
            if status == 200:
                self.send_header("Content-Type", "application/json")
            self.send_header("Content-Length", str(len(out)))
            self.end_headers()
            self.wfile.write(out)
            if err and getattr(self.server, "verbose", False):
                self.log_message("%d %s", status, err)

        do_GET = do_POST = do_PUT = do_DELETE = do_PATCH = _serve

        def log_message(self, fmt, *args):
            if getattr(self.server, "verbose", False):
                super().log_message(fmt, *args)

    return Handler


def serve(api: WireApi, host="127.0.0.1", port=0, verbose=False):
    """Starts a threaded HTTP server in the background; returns (httpd, thread).
    port 0 picks a free port (httpd.server_address[1])."""
    httpd = ThreadingHTTPServer((host, port), make_handler(api))
    httpd.daemon_threads = True
    httpd.verbose = verbose
    th = threading.Thread(target=httpd.serve_forever, daemon=True)
    th.start()
    return httpd, th


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows-log2", type=int, default=12)
    ap.add_argument("--row-bytes", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--host", default="127.0.0.1")
    ap.add_argument("--port", type=int, default=8080)
    ap.add_argument("--max-body-bytes", type=int, default=0, help="0 = 64 KiB")
    ap.add_argument("--max-active-keys", type=int, default=0)
    args = ap.parse_args()
    scanner = GpuScanner(0)
    m = GpuPageMatrix(scanner, 1 << args.rows_log2, args.row_bytes)
    m.fill_seeded(args.seed)
    srv = GpuServer(scanner, m, max_active_keys=args.max_active_keys)
    api = WireApi(srv, max_body_bytes=args.max_body_bytes)
    httpd = ThreadingHTTPServer((args.host, args.port), make_handler(api))
    # Handler threads are joined by server_close() (block_on_close), so none is
    # still inside the request loop when it is destroyed below.
    httpd.daemon_threads = False
    httpd.block_on_close = True
    httpd.verbose = True
    print(f"serving 2^{args.rows_log2} x {args.row_bytes} B on "
          f"http://{args.host}:{httpd.server_address[1]}", file=sys.stderr, flush=True)
    try:
        httpd.serve_forever()
    except KeyboardInterrupt:
        pass
    finally:
        httpd.server_close()
        srv.close()
        m.close()
        scanner.close()


if __name__ == "__main__":
    main()
