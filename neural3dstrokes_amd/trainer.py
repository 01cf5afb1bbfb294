"""The reference's training step for one stroke field, replayed as a HIP graph.

One step of ``train.py:127-144`` for the stroke-field level: render the batch
(the fused operator through ``StrokeField.render``), the final-level data loss
of ``compute_data_loss`` (``source/loss.py:246-302``: 'mse', 'l1', 'charb'
with ``charb_padding`` or 'huber', mean over rays and channels, times
``data_loss_mult``), ``loss.backward()``, ``clip_gradients``
(``source/utils/training.py:67-77``: global norm, value, ``nan_to_num``),
Adam with the parameter ranges (``clip_params``) at the
``learning_rate_decay`` rate, then ``StrokeField.step_update`` (stroke
insertion and reset, host-side, reading the step's gradients for the
shape-gradient moving average).  The gradients step_update reads are the
unclipped ones (the fused Adam clips inside its update): they differ from the
reference's only when ``grad_max_norm`` / ``grad_max_val`` clip or a
gradient is not finite.

Everything from the render to the optimizer update is recorded once as a HIP
graph and replayed; the learning rate lives on the device so a replay uses the
current one.  ``step_update`` runs between replays (as the reference calls it
after ``optimizer.step()``); when it changes the stroke count the step's
shapes change, and the next step records a new graph for the new count.
Results equal the eager step's (``eager=True``).

Out of scope: the proposal levels' interlevel loss, the error field's own
loss and the other regularisers of ``apply_loss`` (``train.py:193-249``), and
textures; the error field, if given, is only sampled by ``step_update``.
"""
import torch

from .optim import Adam, learning_rate_decay

LOSSES = ('mse', 'l1', 'charb', 'huber')


def data_loss(rgb, target, kind='charb', charb_padding=0.001, huber_delta=1.0):
    """``_get_data_loss`` (loss.py:246-261) averaged as compute_data_loss does."""
    r = rgb - target[..., :3]
    if kind == 'mse':
        v = r * r
    elif kind == 'l1':
        v = r.abs()
    elif kind == 'charb':
        v = torch.sqrt(r * r + charb_padding ** 2)
    elif kind == 'huber':
        a = r.abs()
        v = torch.where(a < huber_delta, 0.5 * r * r, huber_delta * (a - 0.5 * huber_delta))
    else:
        raise KeyError(f'Unknown data loss type {kind}')
    return v.mean()


class StrokeFieldTrainer:
    """Training loop state for a ``StrokeField`` on batches of rays with target colours.

    ``step(batch)`` runs one training step on ``batch`` (a dict with ``origins``,
    ``directions``, ``viewdirs``, ``radii``, ``tdist`` and ``rgb``, all on the
    field's device; the same shapes every step) and returns the loss tensor
    (device, not synchronised).
    """

    def __init__(self, field, max_steps, lr_init=0.01, lr_final=0.003, lr_delay_steps=0, lr_delay_mult=1.0,
                 betas=(0.9, 0.99), eps=1e-15, grad_max_norm=0.0, grad_max_val=0.0, data_loss_type='charb',
                 charb_padding=0.001, data_loss_mult=1.0, bg_color=(1.0, 1.0, 1.0), error_field=None, eager=False,
                 render_kw=None):
        self.field = field
        self.max_steps = int(max_steps)
        self.lr_args = (lr_init, lr_final, self.max_steps, lr_delay_steps, lr_delay_mult)
        dev = field.shape_params.device
        self.lr = torch.full((1,), float(lr_init), device=dev)
        groups = [dict(params=[field.shape_params], ranges=field.shape_param_ranges),
                  dict(params=[field.color_params], ranges=field.color_param_ranges),
                  dict(params=[field.density_params], min_value=0.0)]
        groups = [g for g in groups if g['params'][0].requires_grad]
        self.opt = Adam(groups, lr=self.lr, betas=betas, eps=eps, grad_max_norm=grad_max_norm,
                        grad_max_val=grad_max_val, capturable=True)
        if data_loss_type not in LOSSES:
            raise KeyError(f'Unknown data loss type {data_loss_type}')
        self.loss_kw = dict(kind=data_loss_type, charb_padding=charb_padding)
        self.data_loss_mult = float(data_loss_mult)
        self.bg_color = bg_color
        self.error_field = error_field
        self.eager = eager
        self.render_kw = dict(render_kw or {})
        self.graph = None
        self.graph_key = None
        self.static = None
        self.loss = None
        self.cur_step = 0
        self.recordings = 0

    def _fwd_bwd_update(self, b):
        f = self.field
        # the previous step's gradients stay until here: step_update reads them for the
        # shape-gradient moving average, as the reference zeroes them after it
        for p in (f.shape_params, f.color_params, f.density_params):
            if p.grad is not None:
                p.grad.zero_()
        rgb, _, _ = f.render(b['origins'], b['directions'], b['viewdirs'], b['radii'], b['tdist'],
                             bg_color=self.bg_color, **self.render_kw)
        loss = self.data_loss_mult * data_loss(rgb, b['rgb'], **self.loss_kw)
        loss.backward()
        self.opt.step()
        return loss.detach()

    def _record(self, batch):
        """Record the step for the current stroke count (recording does not run it)."""
        self.static = {k: v.detach().clone() for k, v in batch.items()}
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.static_loss = self._fwd_bwd_update(self.static)
        self.recordings += 1

    def step(self, batch):
        self.cur_step += 1
        self.lr.fill_(learning_rate_decay(self.cur_step, *self.lr_args))
        f = self.field
        key = (f.n_strokes(), f.training, tuple(batch['tdist'].shape))
        first = not self.opt.state
        if self.eager or first:
            # eager: also the first step, which creates the optimizer state outside any graph
            self.loss = self._fwd_bwd_update(batch)
        else:
            if self.graph is None or self.graph_key != key:
                self._record(batch)
                self.graph_key = key
            else:
                for k, v in batch.items():
                    self.static[k].copy_(v)
            self.graph.replay()
            self.loss = self.static_loss
        f.step_update(self.cur_step, self.max_steps, error_field=self.error_field)
        return self.loss
